"""The device-resident main loop with hyperbolic self-gravity (hg_run_gravity).

It must give the same state, bit for bit, as the per-step entry points
(hg_timestep + hg_gravity + hg_hydro_step + hg_phidamp) and as the oracle,
in every sweep form, across output stops of the loop, with the same number
of gravity sub-steps.
"""
import numpy as np
import pytest

from hormone_amd import driver as D
from hormone_amd import problems as P
from hormone_amd.native import Hydro
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def _state(eng, g, prm):
    flds = (D.SNAP_FIELDS_MAG if prm.mag_on else D.SNAP_FIELDS) + (D.F_GRVPHI, D.F_GRVPSI)
    return D.fetch(eng, g, flds)


def _same(a, b):
    return all(np.array_equal(a[f].view(np.uint64), b[f].view(np.uint64)) for f in b)


@pytest.mark.parametrize("mag_on", [False, True], ids=["hydro", "mhd"])
@pytest.mark.parametrize("fused", [0, 1, 2], ids=["separate", "last-fused", "x-fused"])
def test_gpu_device_loop_gravity_equals_per_step_and_oracle(hiplib, mag_on, fused):
    print(f"case polytrope24 device loop mag_on={mag_on} fused={fused}", flush=True)
    steps = 10
    out = {}
    for kind in ("device", "per-step", "oracle"):
        prm, g, st = P.make_problem("polytrope3d", n=24, steps=steps, mag_on=mag_on)
        eng = (Oracle if kind == "oracle" else Hydro)(**D.desc_fields(prm, g))
        if kind != "oracle" and fused:
            eng.set_fused(fused)
        D.load_initial_state(eng, prm, g, st)
        r = D.run(eng, prm, g, device_loop=kind == "device")
        nsub = r.nsub_total if kind == "device" else sum(r.nsub)
        out[kind] = (r.tn, r.time, r.cgrav_old, nsub, _state(eng, g, prm))
    tn, tm, co, ns, s = out["oracle"]
    assert tn == steps and ns >= steps - 1
    for kind in ("device", "per-step"):
        tn2, tm2, co2, ns2, s2 = out[kind]
        assert (tn2, tm2, co2, ns2) == (tn, tm, co, ns), kind
        assert _same(s2, s), kind


def test_gpu_device_loop_gravity_across_output_stops(hiplib):
    """Output every 3 steps: the loop returns at each output, the caller
    calls again with the gravity clock it got back, and the end state is
    the one of an uninterrupted per-step run."""
    print("case polytrope24 device loop with output stops", flush=True)
    prm, g, st = P.make_problem("polytrope3d", n=24, steps=10, mag_on=True, tn_out=3)
    eng = Hydro(**D.desc_fields(prm, g))
    eng.set_fused(2)
    D.load_initial_state(eng, prm, g, st)
    t_end, dt_out, t_out = prm.times()
    damp = np.exp(-prm.alpha9wave * prm.courant)
    tim, tn, gt, co, stops = 0.0, 0, 0.0, 0.0, []
    while True:
        tim, tn, t_out, reason, gt, co, _ = eng.run_loop_gravity(
            tim, tn, t_out, dt_out, t_end, prm.outstyle, prm.endstyle, prm.tnlim, prm.tn_out,
            1 << 40, damp, gt, co)
        stops.append((reason, tn))
        if reason != 2:
            break
    assert stops == [(2, 3), (2, 6), (2, 9), (1, 10)]
    prm2, g2, st2 = P.make_problem("polytrope3d", n=24, steps=10, mag_on=True)
    ref = Hydro(**D.desc_fields(prm2, g2))
    D.load_initial_state(ref, prm2, g2, st2)
    r = D.run(ref, prm2, g2)
    assert (tn, tim, co) == (r.tn, r.time, r.cgrav_old)
    assert _same(_state(eng, g, prm), _state(ref, g2, prm2))


def test_gpu_device_loop_gravity_max_steps_then_resume(hiplib):
    """max_steps splits the run; resuming with the returned clocks gives the
    uninterrupted result."""
    print("case polytrope24 device loop max_steps", flush=True)
    prm, g, st = P.make_problem("polytrope3d", n=24, steps=9)
    eng = Hydro(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    t_end, dt_out, t_out = prm.times()
    tim, tn, gt, co = 0.0, 0, 0.0, 0.0
    for chunk in (1, 4, 100):
        tim, tn, t_out, reason, gt, co, _ = eng.run_loop_gravity(
            tim, tn, t_out, dt_out, t_end, prm.outstyle, prm.endstyle, prm.tnlim, prm.tn_out,
            chunk, 0.0, gt, co)
    assert reason == 1 and tn == 9
    prm2, g2, st2 = P.make_problem("polytrope3d", n=24, steps=9)
    ref = Oracle(**D.desc_fields(prm2, g2))
    D.load_initial_state(ref, prm2, g2, st2)
    r = D.run(ref, prm2, g2)
    assert (tn, tim, co) == (r.tn, r.time, r.cgrav_old)
    assert _same(_state(eng, g, prm), _state(ref, g2, prm2))


def test_gpu_two_stage_gravity_substep_closed_form(hiplib):
    """One grktype-2 sub-step on the GPU equals the exact two-stage result
    (tests/test_gravity.py derives it) to 1e-13."""
    from tests.test_gravity import two_stage_substep_errors
    print("case gravity two-stage closed form", flush=True)
    ephi, epsi, edge, term = two_stage_substep_errors(Hydro)
    assert ephi < 1e-13 and epsi < 1e-13, (ephi, epsi)
    assert edge > 1e-9 and term > 1e-4
