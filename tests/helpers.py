"""Shared helpers for the parity tests."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

from hormone_amd import driver as D
from hormone_amd import problems as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("d", "v1", "v2", "v3", "e", "b1", "b2", "b3", "phi")


def meta():
    with open(os.path.join(GOLDEN, "golden_meta.json")) as f:
        return json.load(f)


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def run_problem(engine_cls, name, flat=False, max_steps=None, **kw):
    prm, g, st = P.make_problem(name, flux_limiter="flat" if flat else "modified_mc", **kw)
    eng = engine_cls(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    res = D.run(eng, prm, g, max_steps=max_steps)
    fields = D.SNAP_FIELDS_MAG if prm.mag_on else D.SNAP_FIELDS
    out = D.fetch(eng, g, fields)
    state = {NAMES[i]: out[f] for i, f in enumerate(fields)}
    if prm.compswitch == 2:
        spc = D.fetch(eng, g, tuple(D.F_SPC + q for q in range(prm.spn)))
        for q in range(prm.spn):
            state[f"spc{q + 1}"] = spc[D.F_SPC + q]
    return res, state, eng


def compare_to_meta(res, state, entry):
    """Bitwise comparison against a digest entry of golden_meta.json."""
    assert res.tn == entry["tn"], (res.tn, entry["tn"])
    assert res.time == entry["time"], (res.time, entry["time"])
    bad = [k for k, v in entry["fields"].items() if digest(state[k]) != v["sha256"]]
    assert not bad, f"fields differing from the reference: {bad}"


def compare_to_npz(res, state, name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    assert res.tn == int(z["tn"])
    assert res.time == float(z["time"])
    for k in state:
        a = state[k].reshape(-1)
        b = z[k]
        nbad = int(np.sum(a.view(np.uint64) != b.view(np.uint64)))
        assert nbad == 0, f"{name}: field {k} differs from the reference in {nbad} cells"


def rel_l1(a, b):
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / den) if den > 0 else float(np.abs(a - b).sum())


# Reference runs of features the committed goldens do not exercise
# (tests/golden/make_fixtures.py, EXTRA): name -> (problem, keywords).
# bc tuples are (bc1is bc1os bc2is bc2os bc3is bc3os bc1iv bc1ov bc2iv bc2ov bc3iv bc3ov).
def _bc(axis, code, default=3):
    bc = [default] * 12
    for q in (2 * (axis - 1), 2 * (axis - 1) + 1, 6 + 2 * (axis - 1), 6 + 2 * (axis - 1) + 1):
        bc[q] = code
    return tuple(bc)


EXTRA_CASES = {
    "sod_x_reflective": ("sodshock_x", dict(n=256, bc=_bc(1, 1), t_end=400.0, dt_out=400.0)),
    "sod_x_outgoing": ("sodshock_x", dict(n=256, bc=_bc(1, 2), t_end=300.0, dt_out=300.0)),
    "sod_y_outgoing": ("sodshock_y", dict(n=256, bc=_bc(2, 2), t_end=300.0, dt_out=300.0)),
    "briowu_x_reflective_5steps": ("briowushock_x", dict(n=256, bc=_bc(1, 1), outstyle=2,
                                                         endstyle=2, tnlim=5, tn_out=5)),
    "briowu_z_outgoing": ("briowushock_z", dict(n=256, bc=_bc(3, 2), t_end=300.0, dt_out=300.0)),
    "sod_x_minmod_rk2": ("sodshock_x", dict(n=256, flux_limiter="minmod", rktype=2)),
    "briowu_y_minmod": ("briowushock_y", dict(n=256, flux_limiter="minmod")),
    "sod_x_flat_rk1": ("sodshock_x", dict(n=256, flux_limiter="flat", rktype=1)),
    "ot64_minmod_rk2": ("orszagtang_xy", dict(n=64, flux_limiter="minmod", rktype=2, outstyle=2,
                                              endstyle=2, tnlim=100, tn_out=100)),
    # linear boundary codes 4/5 (make_fixtures.py, LINEAR)
    "sod_x_linear": ("sodshock_x", dict(n=256, bc=(2, 4, 3, 3, 3, 3, 2, 4, 3, 3, 3, 3),
                                        t_end=300.0, dt_out=300.0)),
    "briowu_x_linear5": ("briowushock_x", dict(n=256, bc=(3, 5, 3, 3, 3, 3, 3, 5, 3, 3, 3, 3),
                                               t_end=300.0, dt_out=300.0)),
    "briowu_z_linear5": ("briowushock_z", dict(n=256, bc=(3, 3, 3, 3, 5, 4, 3, 3, 3, 3, 5, 4),
                                               outstyle=2, endstyle=2, tnlim=450, tn_out=450)),
    "ot_yz_linear": ("orszagtang_yz", dict(n=64, bc=(0, 0, 0, 0, 4, 4, 0, 0, 0, 0, 4, 5),
                                           outstyle=2, endstyle=2, tnlim=40, tn_out=40)),
}
LINEAR_CASES = ("sod_x_linear", "briowu_x_linear5", "briowu_z_linear5", "ot_yz_linear")


def run_extra(engine_cls, case, max_steps=None):
    name, kw = EXTRA_CASES[case]
    prm, g, st = P.make_problem(name, **kw)
    eng = engine_cls(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    res = D.run(eng, prm, g, max_steps=max_steps)
    fields = D.SNAP_FIELDS_MAG if prm.mag_on else D.SNAP_FIELDS
    out = D.fetch(eng, g, fields)
    return res, {NAMES[i]: out[f] for i, f in enumerate(fields)}, eng
