"""The Fortran host (hormone_amd/fortran/hg_driver.f90) drives the C ABI
through iso_c_binding and reproduces the reference's golden snapshots.

The parameter files use the reference's namelist groups and values
(para/parameters_base + para/parameters_<simtype>); the driver writes the
final state in the reference's snapshot layout.
"""
import os
import subprocess

import numpy as np
import pytest

from hormone_amd.build import DRIVER
from hormone_amd.snapshot import read_bin
from tests.helpers import GOLDEN, digest, meta

pytestmark = pytest.mark.gpu

PARAMS = {
    "sodshock_x": """&gridcon is=1 ie=500 js=1 je=1 ks=1 ke=1 /
&out_con outstyle=1 endstyle=1 tnlim=100000 t_end=200d0 dt_out=200d0 dt_unit='ms' /
&eos_con eostype=0 gamma=1.4d0 /
&simucon crdnt=0 courant=0.9d0 rktype=3 mag_on=.false. flux_limiter='modified_mc' /
&bouncon bc1is=3 bc1os=3 bc1iv=3 bc1ov=3 /
""",
    "briowushock_z": """&gridcon is=1 ie=1 js=1 je=1 ks=1 ke=500 /
&out_con outstyle=1 endstyle=1 tnlim=100000 t_end=200d0 dt_out=200d0 dt_unit='ms' /
&eos_con eostype=0 gamma=2.0d0 /
&simucon crdnt=0 mag_on=.true. /
&bouncon bc3is=3 bc3os=3 bc3iv=3 bc3ov=3 /
""",
    "orszagtang_xy": """&gridcon is=1 ie=128 js=1 je=128 ks=1 ke=1 /
&out_con outstyle=1 endstyle=1 tnlim=100000 t_end=500d0 dt_out=500d0 dt_unit='ms' /
&eos_con eostype=0 eoserr=1d-15 /
&simucon crdnt=0 mag_on=.true. /
&bouncon bc1is=0 bc1os=0 bc2is=0 bc2os=0 bc3is=0 bc3os=0
         bc1iv=0 bc1ov=0 bc2iv=0 bc2ov=0 bc3iv=0 bc3ov=0 /
""",
}
SHAPES = {"sodshock_x": ((1, 1, 500), False), "briowushock_z": ((500, 1, 1), True),
          "orszagtang_xy": ((1, 128, 128), True)}


@pytest.mark.parametrize("simtype,fused", [(s, 0) for s in PARAMS] + [
    ("sodshock_x", 2), ("briowushock_z", 1), ("orszagtang_xy", 2)])
def test_fortran_driver_matches_golden(simtype, fused, tmp_path, hiplib):
    """The Fortran host (&gpucon fused: hg_set_fused) reproduces the goldens."""
    assert os.path.exists(DRIVER), "hg_driver not built (needs amdflang at build time)"
    par = tmp_path / "parameters"
    out = tmp_path / "final.bin"
    par.write_text(PARAMS[simtype]
                   + f"&gpucon simtype='{simtype}' outfile='{out}' fused={fused} /\n")
    print(f"case fortran driver {simtype} fused={fused}", flush=True)
    r = subprocess.run([DRIVER, str(par)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    shape, mag = SHAPES[simtype]
    snap = read_bin(str(out), shape, mag)
    names = ("d", "v1", "v2", "v3", "e") + (("b1", "b2", "b3", "phi") if mag else ())
    if simtype == "orszagtang_xy":
        entry = meta()[simtype]
        assert snap["tn"] == entry["tn"] and snap["time"] == entry["time"]
        for nm in names:
            assert digest(snap[nm]) == entry["fields"][nm]["sha256"], nm
    else:
        z = np.load(os.path.join(GOLDEN, simtype + ".npz"))
        assert snap["tn"] == int(z["tn"]) and snap["time"] == float(z["time"])
        for nm in names:
            assert np.array_equal(snap[nm].reshape(-1).view(np.uint64), z[nm].view(np.uint64)), nm


POLY_PARAMS = """&gridcon is=1 ie={n} js=1 je={n} ks=1 ke={n}
         xi1s=-1d0 xi1e=1d0 xi2s=-1d0 xi2e=1d0 xi3s=-1d0 xi3e=1d0 /
&out_con outstyle=2 endstyle=2 tnlim={steps} tn_out={steps} t_end=1d30 dt_out=1d0 dt_unit='s' /
&eos_con eostype=0 gamma=2d0 /
&simucon crdnt=0 courant=0.9d0 rktype=3 mag_on=.false. flux_limiter='modified_mc' /
&bouncon bc1is=3 bc1os=3 bc2is=3 bc2os=3 bc3is=3 bc3os=3
         bc1iv=3 bc1ov=3 bc2iv=3 bc2ov=3 bc3iv=3 bc3ov=3 /
&gravcon gravswitch=3 HGfac=5d0 hgcfl=0.9d0 grktype=2 alphagrv=0.1d0 maxtngrv=-1 /
"""


def test_fortran_driver_gravity_matches_python_host(tmp_path, hiplib):
    """The Fortran host drives the hyperbolic self-gravity path (hg_timestep ->
    hg_gravity -> hg_hydro_step) through iso_c_binding and ends bit for bit
    where the Python host does on the same initial state (an n=1 polytrope,
    24^3, 6 steps)."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    from hormone_amd.native import Hydro
    assert os.path.exists(DRIVER), "hg_driver not built (needs amdflang at build time)"
    n, steps = 24, 6
    prm, g, st = P.make_problem("polytrope3d", n=n, steps=steps)
    init = tmp_path / "init.bin"
    with open(init, "wb") as f:
        for a in (st.d, st.p, st.v[0], st.v[1], st.v[2], st.b[0], st.b[1], st.b[2]):
            f.write(np.ascontiguousarray(a, "<f8").tobytes())
        f.write(np.ascontiguousarray(st.extra["grvphi_padded"], "<f8").tobytes())
    par = tmp_path / "parameters"
    out = tmp_path / "final.bin"
    par.write_text(POLY_PARAMS.format(n=n, steps=steps)
                   + f"&gpucon simtype='polytrope' initfile='{init}' outfile='{out}' "
                     f"is_test=.false. gconst=1d0 /\n")
    print("case fortran driver polytrope gravity", flush=True)
    r = subprocess.run([DRIVER, str(par)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    snap = read_bin(str(out), (n, n, n), False)

    eng = Hydro(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    res = D.run(eng, prm, g)
    mine = D.fetch(eng, g, D.SNAP_FIELDS)
    assert snap["tn"] == res.tn == steps and snap["time"] == res.time
    for f, nm in zip(D.SNAP_FIELDS, ("d", "v1", "v2", "v3", "e")):
        assert np.array_equal(snap[nm].view(np.uint64), mine[f].view(np.uint64)), nm
