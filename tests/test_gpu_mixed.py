"""FP32-mixed precision path: the Riemann solver (fast speeds, star states,
state selection) in FP32, in perturbation form about the left face state;
state, division-free reconstruction, limiter and e_int-guard decisions,
face internal energy, base flux and flux-divergence accumulation in FP64.

Not bit-exact by design.  Two kinds of checks:

* The mixed result is compared with the FP64 result at the same physical
  time, and with the *rounding floor* of the problem: the FP64 run started
  from a density perturbed by 1e-14.  The scheme's limiters and e_int guard
  are switches, so on noisy or turbulent data any rounding-level difference
  grows to a problem-dependent floor; a mixed path whose error sits at that
  floor is as close to FP64 as FP64 is to itself under a change of rounding.
* Tolerances measured on smooth early-time flows, and mass conservation to
  FP64 rounding (the flux differences are accumulated in FP64).
"""
import numpy as np
import pytest

import bench as B
from hormone_amd import driver as D
from hormone_amd import problems as P
from hormone_amd.native import Hydro

pytestmark = pytest.mark.gpu


def _start(name, prec, kw, perturb=0.0):
    prm, g, st = P.make_problem(name, precision=prec, **kw)
    if perturb:
        st.d = st.d * (1.0 + perturb)
    eng = Hydro(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    return prm, g, eng, B.Runner(eng, prm, g)


def _fields(prm):
    return D.SNAP_FIELDS_MAG if prm.mag_on else D.SNAP_FIELDS


def _mixed_vs_fp64(name, steps, kw, floor=False):
    prm, g, eng, r = _start(name, "fp32mix", kw)
    r.advance(steps)
    mine = D.fetch(eng, g, _fields(prm))
    prm6, g6, e6, r6 = _start(name, "fp64", kw)
    B.advance_to(r6, r.tim)
    ref = D.fetch(e6, g6, _fields(prm))
    out = {"mixed": D.rel_l1_errors(mine, ref)}
    if floor:
        prm7, g7, e7, r7 = _start(name, "fp64", kw, perturb=1e-14)
        B.advance_to(r7, r.tim)
        out["floor"] = D.rel_l1_errors(D.fetch(e7, g7, _fields(prm)), ref)
    return out


@pytest.mark.parametrize("name,steps,kw", [
    ("kh3d", 60, dict(n=48, steps=60)),
    ("polytrope3d", 40, dict(n=48, steps=40)),
])
def test_mixed_error_at_rounding_floor(hiplib, name, steps, kw):
    print(f"case mixed floor {name}", flush=True)
    res = _mixed_vs_fp64(name, steps, kw, floor=True)
    print(res, flush=True)
    for q in ("d", "e"):
        assert res["mixed"][q] <= 3.0 * res["floor"][q] + 1e-9, (q, res)


@pytest.mark.parametrize("name,steps,kw,tol", [
    ("orszagtang_xy", 20, dict(n=128), 2e-4),
    ("sodshock_x", 300, dict(n=1024, courant=0.4), 1e-3),
    # linear boundary codes on the z faces (the mixed sweep's own
    # instantiation for codes 4, 5, 9, 10)
    ("orszagtang_yz", 30, dict(n=64, bc=(0, 0, 0, 0, 4, 4, 0, 0, 0, 0, 4, 5), outstyle=2,
                               endstyle=2, tnlim=40, tn_out=40), 1e-3),
])
def test_mixed_close_to_fp64(hiplib, name, steps, kw, tol):
    print(f"case mixed {name}", flush=True)
    res = _mixed_vs_fp64(name, steps, kw)
    print(res, flush=True)
    for q in ("d", "e"):
        assert res["mixed"][q] < tol, (q, res)


def test_mixed_conserves_mass_to_fp64_rounding(hiplib):
    prm, g, eng, r = _start("orszagtang_xy", "fp32mix", dict(n=128))
    r.advance(40)
    s = D.fetch(eng, g, [D.F_D])
    prm0, g0, st = P.make_problem("orszagtang_xy", n=128)
    m0 = float(st.d.sum())
    m1 = float(s[D.F_D].sum())
    assert abs(m1 - m0) / m0 < 1e-12


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("name,kw,steps", [("kh3d", dict(n=28, steps=8), 8),
                                           ("orszagtang_xy", dict(n=90), 12)])
def test_mixed_fused_equals_mixed_separate(hiplib, name, kw, steps, mode):
    """The FP32-mixed sweep fused with the (FP64) RK update gives the
    separate mixed launches' result bit for bit."""
    print(f"case mixed fused {name} mode={mode}", flush=True)
    out = []
    for fused in (0, mode):
        prm, g, eng, r = _start(name, "fp32mix", kw)
        eng.set_fused(fused)
        r.advance(steps)
        out.append((r.tn, r.tim, D.fetch(eng, g, _fields(prm))))
    (n1, t1, a), (n2, t2, b) = out
    assert (n1, t1) == (n2, t2)
    for f in a:
        assert np.array_equal(a[f].view(np.uint64), b[f].view(np.uint64)), f
