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
    r6.advance_to(r.tim)
    ref = D.fetch(e6, g6, _fields(prm))
    out = {"mixed": D.rel_l1_errors(mine, ref)}
    if floor:
        prm7, g7, e7, r7 = _start(name, "fp64", kw, perturb=1e-14)
        r7.advance_to(r.tim)
        out["floor"] = D.rel_l1_errors(D.fetch(e7, g7, _fields(prm)), ref)
    return out


# The target: relative L1 <= 1e-5 per variable against FP64.
TARGET = 1e-5
# phi (the GLM cleaning field) is noise-level; its relative L1 is not a
# meaningful accuracy figure and is not asserted.
CHECKED = ("d", "v1", "v2", "v3", "e", "b1", "b2", "b3")


@pytest.mark.parametrize("name,steps,kw", [
    ("kh3d", 60, dict(n=48, steps=60)),
    ("polytrope3d", 40, dict(n=48, steps=40)),
    ("polytrope3d", 40, dict(n=48, steps=40, mag_on=True)),
    ("sodshock_x", 300, dict(n=1024, courant=0.4)),
])
def test_mixed_error_at_rounding_floor(hiplib, name, steps, kw):
    """Every variable: the mixed error is within 3x what FP64 itself shows
    under a 1e-14 density perturbation, and within the 1e-5 target wherever
    that FP64 floor is below it.  Measured (mixed / floor): KH 48^3 60
    steps d 1.9e-6/2.0e-6, v 4.0-7.9e-6/3.9-7.7e-6, e 3.3e-6/3.3e-6, B
    4.3e-6-1.06e-5/4.2e-6-1.05e-5; polytrope 48^3 40 steps d 3.3e-4/1.7e-4,
    v 2.4e-2/2.0e-2, e 1.9e-4/9.3e-5; Sod 1024 300 steps d 9.2e-5/1.1e-4,
    v1 7.7e-4/1.9e-3, e 2.9e-5/7.1e-5."""
    print(f"case mixed floor {name} {kw}", flush=True)
    res = _mixed_vs_fp64(name, steps, kw, floor=True)
    print(res, flush=True)
    for q in CHECKED:
        if q not in res["mixed"]:
            continue
        m, f = res["mixed"][q], res["floor"][q]
        assert m <= 3.0 * f + 1e-12, (q, res)
        if f < TARGET:
            assert m <= TARGET, (q, res)


@pytest.mark.parametrize("name,steps,kw,tol", [
    # smooth early-time flow: measured 1.4e-10 .. 1.9e-10 per variable
    ("orszagtang_xy", 20, dict(n=128), 1e-9),
    # linear boundary codes on the z faces (the mixed sweep's own
    # instantiation for codes 4, 5, 9, 10): measured 3.3e-10 .. 5.1e-10
    ("orszagtang_yz", 30, dict(n=64, bc=(0, 0, 0, 0, 4, 4, 0, 0, 0, 0, 4, 5), outstyle=2,
                               endstyle=2, tnlim=40, tn_out=40), 3e-9),
])
def test_mixed_close_to_fp64(hiplib, name, steps, kw, tol):
    """Smooth flows, before the limiters' switches amplify rounding: every
    variable within `tol` (far inside the 1e-5 target)."""
    print(f"case mixed {name}", flush=True)
    res = _mixed_vs_fp64(name, steps, kw)
    print(res, flush=True)
    for q in CHECKED:
        if q in res["mixed"]:
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
