"""Snapshots and restart (the reference's write_bin / readbin / restart,
output.f90:697-795, readbin.f90:16-115, restart.f90:15-64), pinned by a
reference run of Orszag-Tang 64^2 that wrote a snapshot at step 50 and by the
reference's own restart from that snapshot to step 100."""
import hashlib
import os

import numpy as np
import pytest

from hormone_amd import driver as D
from hormone_amd import problems as P
from hormone_amd.snapshot import read_bin, write_bin
from tests.helpers import GOLDEN, NAMES, compare_to_meta, meta

RESTART = dict(n=64, outstyle=2, endstyle=2, tnlim=100, tn_out=50)


def step50():
    z = np.load(os.path.join(GOLDEN, "ot64_step50.npz"))
    snap = {k: z[k].reshape(1, 64, 64) for k in ("d", "v1", "v2", "v3", "e", "b1", "b2", "b3",
                                                 "phi")}
    snap["tn"], snap["time"] = int(z["tn"]), float(z["time"])
    return snap


def test_write_bin_reproduces_the_reference_file():
    """Our writer turns the reference's step-50 state into the very bytes of
    the reference's own snapshot file; the reader reads them back."""
    snap = step50()
    raw = write_bin(snap, mag_on=True)
    m = meta()["ot64_step50"]
    assert len(raw) == m["file_bytes"]
    assert hashlib.sha256(raw).hexdigest() == m["file_sha256"]
    path = "/tmp/hormone_restart_test.bin"
    with open(path, "wb") as f:
        f.write(raw)
    back = read_bin(path, (1, 64, 64), True)
    os.remove(path)
    assert back["tn"] == 50 and back["time"] == snap["time"]
    for k in ("d", "e", "b2", "phi"):
        assert np.array_equal(back[k], snap[k])


def restart_run(engine_cls):
    prm, g, _ = P.make_problem("orszagtang_xy", **RESTART)
    eng = engine_cls(**D.desc_fields(prm, g))
    start = D.load_restart_state(eng, prm, g, step50())
    res = D.run(eng, prm, g, start=start)
    out = D.fetch(eng, g, D.SNAP_FIELDS_MAG)
    return res, {NAMES[i]: out[f] for i, f in enumerate(D.SNAP_FIELDS_MAG)}


def test_oracle_restart_matches_the_reference_restart():
    """Restart from the snapshot (conserve from the primitives read, e_int
    by pressure) and run to step 100: the reference's own restarted run,
    bit for bit (it differs from the uninterrupted run, which the reference
    does not reproduce on restart either)."""
    from oracle.pyoracle import Oracle
    res, state = restart_run(Oracle)
    compare_to_meta(res, state, meta()["ot64_restart50_step100"])
    assert meta()["ot64_restart50_step100"]["time"] != meta()["ot64_step100"]["time"]


@pytest.mark.gpu
def test_gpu_restart_matches_the_reference_restart(hiplib):
    from hormone_amd.native import Hydro
    res, state = restart_run(Hydro)
    compare_to_meta(res, state, meta()["ot64_restart50_step100"])


@pytest.mark.gpu
def test_gpu_uninterrupted_ot64_matches_the_reference_step100(hiplib):
    from hormone_amd.native import Hydro
    from tests.helpers import run_problem
    res, state, _ = run_problem(Hydro, "orszagtang_xy", **RESTART)
    compare_to_meta(res, state, meta()["ot64_step100"])


def test_grid_file_matches_the_reference_file():
    """gridfile.bin (write_grid_bin, output.f90:432-452) for Orszag-Tang
    64^2: the reference's own file, element for element, except dx(is-2) of
    each axis, which the reference never sets (gridset.f90:54-59 starts at
    is-1) and so writes whatever that memory held."""
    from hormone_amd.snapshot import read_grid_bin, write_grid_bin
    prm, g, _ = P.make_problem("orszagtang_xy", n=64)
    raw = write_grid_bin(g)
    ref = open(os.path.join(GOLDEN, "gridfile_ot64.bin"), "rb").read()
    assert len(raw) == len(ref) == meta()["gridfile_ot64"]["file_bytes"]
    assert raw[:4] == ref[:4] and raw[-4:] == ref[-4:]  # record length markers
    a, b = read_grid_bin(raw, (64, 64, 1)), read_grid_bin(ref, (64, 64, 1))
    for k in a:
        lo = 1 if k.startswith("dx") and not k.startswith("dxi") else 0
        assert np.array_equal(a[k][lo:], b[k][lo:]), k


def _gravity_run(engine_cls, steps, snap=None):
    """Polytrope 16^3 with hyperbolic gravity: `steps` steps from the start,
    or from snapshot `snap` to step `steps`."""
    prm, g, st = P.make_problem("polytrope3d", n=16, steps=steps)
    eng = engine_cls(**D.desc_fields(prm, g))
    if snap is None:
        D.load_initial_state(eng, prm, g, st)
        res = D.run(eng, prm, g)
    else:
        start = D.load_restart_state(eng, prm, g, snap, st.extra["grvphi_padded"])
        res = D.run(eng, prm, g, start=start)
    return prm, g, eng, res


def test_gravity_snapshot_round_trip_and_restart():
    """Snapshots with the gravity groups (output.f90:735-746: grvphi, then
    grvpsi and cgrav_old) read back exactly, and a restart from one (the
    potential, psi and cgrav_old restored, grvtime = time, restart.f90:59)
    on a fresh engine continues bit for bit like the engine that wrote the
    snapshot restarted from it in place: the snapshot carries all the
    state the step needs.  Dropping psi or cgrav_old changes the result.
    (An uninterrupted run is not the comparison: the restart rebuilds the
    momenta from the primitives, an ulp off, and grvtime = time where the
    running clock trails by an ulp; the polytrope's limiter and e_int-guard
    branches turn that into differences of 1e-4, as in the reference's
    own restart.)"""
    from oracle.pyoracle import Oracle
    from hormone_amd.snapshot import read_bin
    prm, g, eng, res = _gravity_run(Oracle, 6)
    snap = D.snapshot(eng, g, prm, res)
    raw = write_bin(snap, mag_on=False, gravswitch=3)
    n = g.ncells
    assert len(raw) == 4 * 4 + 4 + 8 + 5 * 8 * n + 4 * 4 + 2 * 8 * n + 8
    path = "/tmp/hormone_grav_snapshot.bin"
    with open(path, "wb") as f:
        f.write(raw)
    back = read_bin(path, g.shape, False, gravswitch=3)
    os.remove(path)
    assert back["cgrav_old"] == snap["cgrav_old"] and back["cgrav_old"] > 0.0
    for k in ("d", "e", "grvphi", "grvpsi"):
        assert np.array_equal(back[k], snap[k]), k
    assert write_bin(back, mag_on=False, gravswitch=3) == raw
    # the engine that wrote the snapshot, restarted from it in place
    prm12, g12, st12 = P.make_problem("polytrope3d", n=16, steps=12)
    start = D.load_restart_state(eng, prm12, g12, back, st12.extra["grvphi_padded"])
    D.run(eng, prm12, g12, start=start)
    cont = D.fetch(eng, g, (D.F_D, D.F_E, D.F_GRVPHI, D.F_GRVPSI))
    _, _, e_rest, r_rest = _gravity_run(Oracle, 12, back)
    assert r_rest.tn == 12
    rest = D.fetch(e_rest, g, tuple(cont))
    for f in cont:
        assert np.array_equal(rest[f].view(np.uint64), cont[f].view(np.uint64)), f
    for k, v in (("grvpsi", np.zeros_like(back["grvpsi"])), ("cgrav_old", 0.5 * back["cgrav_old"])):
        _, _, e_bad, _ = _gravity_run(Oracle, 12, dict(back, **{k: v}))
        assert not np.array_equal(D.fetch(e_bad, g, (D.F_GRVPHI,))[D.F_GRVPHI],
                                  cont[D.F_GRVPHI]), k


@pytest.mark.gpu
def test_gpu_gravity_snapshot_and_restart_match_oracle(hiplib):
    """The GPU's snapshot of a gravity run (potential, psi, cgrav_old) is the
    oracle's byte for byte, and restarting from it gives the oracle's
    restarted run bit for bit."""
    from hormone_amd.native import Hydro
    from oracle.pyoracle import Oracle
    print("case gravity snapshot/restart polytrope16", flush=True)
    raws, states = [], []
    for cls in (Hydro, Oracle):
        prm, g, eng, res = _gravity_run(cls, 6)
        raws.append(write_bin(D.snapshot(eng, g, prm, res), mag_on=False, gravswitch=3))
    assert raws[0] == raws[1]
    path = "/tmp/hormone_grav_snapshot_gpu.bin"
    with open(path, "wb") as f:
        f.write(raws[0])
    back = read_bin(path, g.shape, False, gravswitch=3)
    os.remove(path)
    for cls in (Hydro, Oracle):
        _, _, e, r = _gravity_run(cls, 12, back)
        states.append(D.fetch(e, g, (D.F_D, D.F_E, D.F_V3, D.F_GRVPHI, D.F_GRVPSI)))
    for f in states[1]:
        assert np.array_equal(states[0][f].view(np.uint64), states[1][f].view(np.uint64)), f
