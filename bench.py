#!/usr/bin/env python3
"""Zone-updates/s of HORMONE's explicit hydro step on one MI355X.

Workload (the configuration the headline metric is quoted on): 2D
Orszag-Tang MHD vortex, 512^2 periodic on [0,1]^2, gamma 5/3, MUSCL
(modified MC) + HLLD with GLM cleaning, SSP-RK3, courant 0.9, FP64, run to
t = 0.5.  Synthetic initial data (the closed-form vortex); state resident in
HBM.  A step is one full timestep (CFL reduction + 3 RK stages + phidamp).

Timed region: K steps after W warm-up steps, through the device-resident
main loop (hg_run).  After it the run continues, untimed, to step 2200, where
the state is compared bit for bit with the reference program's own run
(tests/golden/golden_meta.json, entry ot512_2200steps); then on until the
reference's own FP64 run stops (negative internal energy during step 2250,
before t = 0.5), which must happen at the same step and cell.

The roofline object is for the dominant kernel (the fused direction sweep),
timed with HIP events on the library's stream inside the timed region; its
algorithmic bytes per launch are dim*(nw+nv)*8 per cell (primitives read
once, one flux divergence written per direction).  cpu_baseline is the CPU
oracle (a C/OpenMP restatement of the reference) on this host's cores over a
bounded prefix of the same run.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    return ap.parse_args()


def cpu_baseline(n: int, seconds: float):
    """Oracle (C/OpenMP restatement of the reference) on the host cores."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    from oracle import pyoracle

    prm, g, st = P.make_problem("orszagtang_xy", n=n)
    eng = pyoracle.Oracle(**D.desc_fields(prm, g))
    pyoracle.set_threads(0)
    cores = pyoracle.lib().or_get_threads()
    D.load_initial_state(eng, prm, g, st)
    t0 = time.perf_counter()
    D.run(eng, prm, g, max_steps=2)
    per = (time.perf_counter() - t0) / 2
    steps = max(2, min(200, int(seconds / max(per, 1e-6))))
    eng2 = pyoracle.Oracle(**D.desc_fields(prm, g))
    D.load_initial_state(eng2, prm, g, st)
    r = D.run(eng2, prm, g, max_steps=steps)
    zus = r.tn * g.ncells / r.wall_s
    return {"value": zus, "unit": "zone-updates/s", "cores": cores, "kind": "port",
            "sample": f"first {r.tn} steps of the same {n}^2 Orszag-Tang run "
                      f"({r.wall_s:.1f} s, {cores} OpenMP threads)"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              f"(the sharded path is a design only).", file=sys.stderr)
        sys.exit(2)
    import torch

    from hormone_amd import driver as D
    from hormone_amd import native
    from hormone_amd import problems as P

    prm, g, st = P.make_problem("orszagtang_xy", n=a.n)
    eng = native.Hydro(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    t_end, dt_out, t_out = prm.times()
    import math
    damp = math.exp(-prm.alpha9wave * prm.courant)
    tim, tn = 0.0, 0

    def advance(k):
        nonlocal tim, tn, t_out
        left = k
        while left > 0:
            n0 = tn
            try:
                tim, tn, t_out, reason = eng.run_loop(tim, tn, t_out, dt_out, t_end,
                                                      prm.outstyle, prm.endstyle, prm.tnlim,
                                                      prm.tn_out, left, damp)
            except native.HgError as err:
                tn, tim = err.tn, err.time
                raise
            left -= tn - n0
            if reason == native.STOP_END:
                return False
        return True

    advance(a.warmup)
    if not a.no_profile:
        eng.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ok = advance(a.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    steps_done = a.steps if ok else None
    assert ok, "the run ended inside the timed region; lower --steps"
    zus = steps_done * g.ncells / wall
    roof = None
    if not a.no_profile:
        ms, nl, bpl = eng.profile_read(native.PROF_SWEEP)
        rk_ms, rk_n, _ = eng.profile_read(native.PROF_RK)
        eng.profile(False)
        avg_s = ms / 1e3 / nl
        ach = bpl / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_sweep (x+y directions, one launch per RK stage)",
                "avg_launch_us": round(avg_s * 1e6, 2), "launches": nl,
                "algorithmic_bytes_per_launch": bpl,
                "share_of_step": round(ms / (wall * 1e3), 3),
                "rk_avg_launch_us": round(rk_ms / max(rk_n, 1) * 1e3, 2)}
    parity = None
    if not a.no_check and a.n == 512:
        # Untimed: continue to step 2200 and compare with the reference
        # program's own run, then run on until the reference's stop (its
        # FP64 run ends with negative internal energy before t = 0.5).
        import hashlib
        import numpy as np
        from tests.helpers import meta
        m = meta()
        advance(2200 - tn)
        out = D.fetch(eng, g, D.SNAP_FIELDS_MAG)
        names = ("d", "v1", "v2", "v3", "e", "b1", "b2", "b3", "phi")
        e = m.get("ot512_2200steps")
        same = e is not None and tn == e["tn"] and tim == e["time"] and all(
            hashlib.sha256(np.ascontiguousarray(out[f], "<f8").tobytes()).hexdigest()
            == e["fields"][nm]["sha256"] for f, nm in zip(D.SNAP_FIELDS_MAG, names))
        parity = {"step": tn, "bitwise_equal_to_reference_run_at_step_2200": bool(same)}
        fail = m.get("ot512_t05_failure")
        try:
            advance(1 << 30)
            parity["reference_stop_reproduced"] = False
        except native.HgError as err:
            parity["stop"] = {"code": err.code, "completed_steps": err.tn,
                              "cell": list(err.cell)}
            parity["reference_stop_reproduced"] = bool(
                fail and err.code == native.HG_ERR_NEG_EINT
                and err.tn == fail["completed_steps"] and list(err.cell) == fail["cell"])
    cpu = None if a.no_cpu else cpu_baseline(a.n, a.cpu_seconds)
    line = {
        "metric": "zone-updates/s (FP64, Orszag-Tang 512^2 MHD)",
        "value": zus, "unit": "zone-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": wall / steps_done * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (closed-form Orszag-Tang initial state)",
        "config": {"workload": f"orszagtang_xy {a.n}x{a.n} MHD+GLM, MUSCL+HLLD, RK3, "
                               f"courant 0.9, FP64 (run toward t=0.5)", "cells": g.ncells,
                   "precision": "fp64"},
        "roofline": roof, "cpu_baseline": cpu, "parity": parity,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
