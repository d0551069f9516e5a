#!/usr/bin/env python3
"""Zone-updates/s of HORMONE's explicit hydro step on one MI355X.

Default workload (the configuration the headline metric is quoted on):
2D Orszag-Tang MHD vortex, 512^2 periodic on [0,1]^2, gamma 5/3, MUSCL
(modified MC) + HLLD with GLM cleaning, SSP-RK3, courant 0.9, FP64.  The
initial data is the closed-form vortex (synthetic); the state stays in HBM.
A step is one full timestep (CFL reduction + 3 RK stages + phidamp), and a
zone-update is one interior cell advanced by one step.

Timed region: K steps after W warm-up steps, through the device-resident
main loop (hg_run), bracketed by torch.cuda.synchronize().  The warm-up
picks the sweep's block size and whether one direction's sweep is fused
with the RK update (none, x or the last direction; --fused) by timing real
steps; every choice gives the same bits.  Tuning steps beyond W are
ordinary steps of the run.  Afterwards the
run continues, untimed, to step 2200, where the state is compared bit for bit
with the reference program's own run (tests/golden/golden_meta.json, entry
ot512_2200steps).  It then runs on until the reference's own FP64 run stops
(negative internal energy during step 2250, before t = 0.5); the stop must
happen at the same step and cell.

Other workloads (--workload):
  polytrope256  3D 256^3 hydro + Cartesian hyperbolic self-gravity, n=1
                polytrope (M=1, R=0.3, G=1, gamma=2) in [-1,1]^3, 200 steps;
                reports mass and energy drift
  kh384         3D 384^3 MHD Kelvin-Helmholtz, periodic, 100 steps
  serving       continuous-batching ensemble: seeded Poisson arrivals (50/s)
                of independent 2D/3D MHD Riemann or blast problems, grids
                log-uniform from 64^2 to 192^3 cells, each to t = 0.1, admitted
                against an HBM budget into one batch of persistent kernels;
                value = aggregate zone-updates/s, with p50/p99 time-to-solution
With --precision fp32mix and --vs-fp64, the same workload is also run in FP64
and the per-variable relative L1 sum|q-q64|dV / sum|q64|dV is reported.

roofline: the dominant kernel is the fused direction sweep.  Its average
launch duration is timed right after the timed region: back-to-back sweep
launches of the same state on the library's stream between two HIP events
(the timed region itself carries no per-dispatch events, which would add a
completion wait to every launch).  Its algorithmic bytes per launch are
dim*(nw+nv)*8 per cell: primitives read once, one flux divergence written per
direction.
cpu_baseline: the CPU oracle (C/OpenMP restatement of the reference) on this
host's cores over a bounded prefix of the same run.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
import sys
import time

import numpy as np

# CPU baseline threads: one per physical core, packed (read by the OpenMP
# runtime when it starts, so set before anything loads it)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NAMES = ("d", "v1", "v2", "v3", "e", "b1", "b2", "b3", "phi")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="ot512",
                    choices=["ot512", "polytrope256", "kh384", "serving"])
    ap.add_argument("--requests", type=int, default=64, help="serving: number of requests")
    ap.add_argument("--hbm-budget-gib", type=float, default=32.0,
                    help="serving: HBM admission budget")
    ap.add_argument("--n", type=int, default=None, help="cells per side (default per workload)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32mix"])
    ap.add_argument("--mag", action="store_true", help="polytrope: MHD variant with B=0")
    ap.add_argument("--vs-fp64", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--serving-lanes", default="50000,500000",
                    help="serving: cell-count bounds of the concurrently served size classes "
                         "(comma separated; empty = one batch)")
    ap.add_argument("--sweep-waves", type=int, default=0,
                    help="sweep wavefronts per block (0: autotune during warm-up)")
    ap.add_argument("--fused", default="auto", choices=["auto", "last", "x", "off"],
                    help="a sweep direction fused with the RK update: the last solved one, "
                         "x, or none (auto: time each during warm-up, keep the fastest)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 256 MiB (the reference kernel for "
                         "scaling rocprof FETCH_SIZE/WRITE_SIZE to bytes)")
    return ap.parse_args()


def copy_calibration():
    """Five device-to-device copies of 256 MiB each: a kernel whose HBM
    traffic is known (256 MiB read + 256 MiB written per launch), profiled
    with the same counters as the hydro kernels to scale them to bytes."""
    import torch
    n = 256 * 2 ** 20 // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"launches": 5, "read_bytes": n * 8,
                                           "write_bytes": n * 8}}))


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "pmc_traffic.json")


def measured_traffic(workload, precision, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprof counter
    runs (FETCH_SIZE + WRITE_SIZE scaled by the copy calibration), or None
    when no measurement of this workload exists."""
    try:
        with open(TRAFFIC_FILE) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None
    e = tab.get("entries", {}).get(f"{workload}/{precision}/{kernel}")
    return None if e is None else e.get("bytes_per_launch")


# steps the warm-up may spend on launch tuning beyond --warmup (the run's
# step limit leaves room for them)
TUNE_STEPS = 64


def setup(a, precision):
    from hormone_amd import problems as P
    limit = a.warmup + a.steps + TUNE_STEPS
    if a.workload == "ot512":
        return P.make_problem("orszagtang_xy", n=a.n or 512, precision=precision)
    if a.workload == "polytrope256":
        return P.make_problem("polytrope3d", n=a.n or 256, steps=limit, mag_on=a.mag,
                              precision=precision)
    return P.make_problem("kh3d", n=a.n or 384, steps=limit, precision=precision)


def cpu_baseline(a, seconds: float):
    """Oracle (C/OpenMP restatement of the reference) on the host cores."""
    from hormone_amd import driver as D
    from oracle import pyoracle

    prm, g, st = setup(a, "fp64")
    pyoracle.set_threads(0)
    cores = pyoracle.lib().or_get_threads()
    eng = pyoracle.Oracle(**D.desc_fields(prm, g))
    pyoracle.set_threads(0)
    D.load_initial_state(eng, prm, g, st)
    t0 = time.perf_counter()
    D.run(eng, prm, g, max_steps=2)
    per = (time.perf_counter() - t0) / 2
    steps = max(2, min(3000, int(seconds / max(per, 1e-6))))
    eng2 = pyoracle.Oracle(**D.desc_fields(prm, g))
    pyoracle.set_threads(0)
    D.load_initial_state(eng2, prm, g, st)
    r = D.run(eng2, prm, g, max_steps=steps)
    zus = r.tn * g.ncells / r.wall_s
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": zus, "unit": "zone-updates/s", "cores": cores, "kind": "port",
            "sample": f"first {r.tn} steps of the same {a.workload} run "
                      f"({r.wall_s:.1f} s, {cores} OpenMP threads, OMP_PROC_BIND=close, "
                      f"OMP_PLACES=cores, {model})"}


class Runner:
    """Drives one engine through the reference's main loop, step by step or
    with the device-resident loop, keeping the clock on the host."""

    def __init__(self, eng, prm, g):
        self.eng, self.prm, self.g = eng, prm, g
        self.t_end, self.dt_out, self.t_out = prm.times()
        self.damp = math.exp(-prm.alpha9wave * prm.courant)
        self.tim, self.tn = 0.0, 0
        self.grvtime, self.cgrav_old = 0.0, 0.0
        self.nsub = []

    def advance(self, k):
        """k more steps; False when the run's end is reached."""
        from hormone_amd import native
        prm, eng = self.prm, self.eng
        if prm.gravswitch == 3:
            for _ in range(k):
                dt, ch, cg, dtg = eng.timestep(self.tim, self.t_out, self.t_end, prm.outstyle,
                                               prm.endstyle, prm.fixed_dt, grav=True)
                if self.tn == 0:
                    self.grvtime = 0.0
                else:
                    self.grvtime, self.cgrav_old, _, ns = eng.gravity(
                        self.tim, dt, self.grvtime, cg, self.cgrav_old, dtg)
                    self.nsub.append(ns)
                eng.hydro_step(dt, ch)
                if prm.mag_on and self.g.dim > 1:
                    eng.phidamp(self.damp)
                self.tim += dt
                self.tn += 1
            return True
        left = k
        while left > 0:
            n0 = self.tn
            try:
                self.tim, self.tn, self.t_out, reason = eng.run_loop(
                    self.tim, self.tn, self.t_out, self.dt_out, self.t_end, prm.outstyle,
                    prm.endstyle, prm.tnlim, prm.tn_out, left, self.damp)
            except native.HgError as err:
                self.tn, self.tim = err.tn, err.time
                raise
            left -= self.tn - n0
            if reason == native.STOP_END:
                return left <= 0
        return True


def advance_to(runner, t_target):
    """Step a Runner until its clock reaches t_target exactly: the end
    condition becomes endstyle 1 (time >= t_end), whose dt clipping
    (timestep.f90:85) lands the last step on t_target; output stops are off."""
    from hormone_amd import native
    prm, eng = runner.prm, runner.eng
    if prm.gravswitch == 3:
        while runner.tim < t_target:
            dt, ch, cg, dtg = eng.timestep(runner.tim, t_target, t_target, 0, 1, 0.0, grav=True)
            if runner.tn == 0:
                runner.grvtime = 0.0
            else:
                runner.grvtime, runner.cgrav_old, _, ns = eng.gravity(
                    runner.tim, dt, runner.grvtime, cg, runner.cgrav_old, dtg)
            eng.hydro_step(dt, ch)
            if prm.mag_on and runner.g.dim > 1:
                eng.phidamp(runner.damp)
            runner.tim += dt
            runner.tn += 1
        return
    while runner.tim < t_target:
        runner.tim, runner.tn, _, reason = eng.run_loop(
            runner.tim, runner.tn, t_target, 0.0, t_target, 0, 1, 1 << 40, 1, 1 << 40,
            runner.damp)
        if reason == native.STOP_END:
            return


def sums(D, eng, g, prm):
    """Mtot, Etot (+ gravitational energy) as output.f90:167-200."""
    f = [D.F_D, D.F_E] + ([D.F_GRVPHI] if prm.gravswitch == 3 else [])
    s = D.fetch(eng, g, f)
    dv = (g.ax[0].dxi[2:-2][None, None, :] * g.ax[1].dxi[2:-2][None, :, None]
          * g.ax[2].dxi[2:-2][:, None, None])
    m = float((s[D.F_D] * dv).sum())
    e = float((s[D.F_E] * dv).sum())
    if prm.gravswitch == 3:
        e += float((s[D.F_D] * 0.5 * s[D.F_GRVPHI] * dv).sum())
    return m, e


def serving_bench(a):
    """The serving ensemble: arrivals honoured in wall-clock time from the
    start, one request = one independent simulation to t = 0.1."""
    import torch

    from hormone_amd import serving as S
    reqs = S.make_requests(a.requests, rate=50.0, seed=7)
    # spot check after the run: the smallest and the largest request, rerun
    # alone, must end in the same state bit for bit as in the mix
    by_size = sorted(reqs, key=lambda r: r.cells)
    check = {by_size[0].rid, by_size[-1].rid}
    torch.cuda.synchronize()
    lanes = tuple(int(x) for x in a.serving_lanes.split(",") if x) if a.serving_lanes else ()
    rep = S.serve(reqs, hbm_budget=a.hbm_budget_gib * 2 ** 30, steps_per_round=4,
                  digest_ids=check, lanes=lanes)
    torch.cuda.synchronize()
    sm = rep.summary()
    done = [r for r in reqs if r.finished is not None and r.error is None]
    same = {}
    for r in reqs:
        if r.rid in check and r.digest is not None:
            tn, tm, dg = S.run_alone(r)
            same[f"request {r.rid} ({r.dim}D, {r.n}^{r.dim})"] = bool(
                dg == r.digest and tn == r.steps)
    line = {
        "metric": "zone-updates/s (serving ensemble, aggregate)", "value": sm["aggregate_zone_updates_per_s"],
        "unit": "zone-updates/s", "n_gpus": 1, "steps": int(sum(r.steps for r in done)),
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded request generator)",
        "config": {"workload": f"serving: {a.requests} Poisson(50/s) MHD Riemann/blast requests, "
                               f"64^2..192^3 cells, t_end 0.1, HBM budget {a.hbm_budget_gib:g} GiB",
                   "requests": a.requests,
                   "size_classes": [f"<= {b} cells" for b in lanes] + ["larger"]},
        "roofline": None, "cpu_baseline": None,
        "p50_time_to_solution_s": sm["p50_s"], "p99_time_to_solution_s": sm["p99_s"],
        "completed": sm["completed"], "failed": sm["failed"], "makespan_s": sm["makespan_s"],
        "zone_updates": rep.zone_updates, "tuned_shape": sm["tuned_shape"],
        "max_in_flight": max((e["active"] for e in rep.events), default=None),
        "alone_equals_in_mix": same,
        "failures": [{"request": r.rid, "cells": r.cells, "kind": r.kind, "error": r.error}
                     for r in reqs if r.error is not None],
    }
    print(json.dumps(line))


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.workload == "serving":
        if a.gpus != 1:
            print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
                  f"(the sharded path is a design only).", file=sys.stderr)
            sys.exit(2)
        serving_bench(a)
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              f"(the sharded path is a design only).", file=sys.stderr)
        sys.exit(2)
    defaults = {"ot512": (2000, 20), "polytrope256": (190, 10), "kh384": (90, 10)}
    if a.steps is None:
        a.steps = defaults[a.workload][0]
        if a.workload == "ot512" and a.precision == "fp32mix":
            # the FP32-mixed run leaves the FP64 trajectory and, like the
            # reference's own run near step 2250, stops on a non-positive face
            # state before step 2020; time an earlier window
            a.steps = 1000
    if a.warmup is None:
        a.warmup = defaults[a.workload][1]
    import torch

    from hormone_amd import driver as D
    from hormone_amd import native

    prm, g, st = setup(a, a.precision)
    eng = native.Hydro(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    run = Runner(eng, prm, g)
    m0, e0 = sums(D, eng, g, prm)

    tuned = None
    if a.sweep_waves:
        eng.set_launch(a.sweep_waves)
        run.advance(a.warmup)
    else:
        # autotune the sweep's block size on warm-up steps (launch shapes do
        # not change results), then finish the warm-up with the winner
        from hormone_amd.batch import tune_single
        per = max(1, a.warmup // 5)
        best, times = tune_single(eng, run.advance, steps=per)
        tuned = {"sweep_block_waves": best, "us_per_launch": {str(k): round(v, 2)
                                                              for k, v in times.items()}}
        left = a.warmup - per * len(times)
        if left > 0:
            run.advance(left)
    fused = False
    names = {0: "separate", 1: "last", 2: "x"}
    if a.fused in ("last", "x"):
        fused = {"last": 1, "x": 2}[a.fused]
        eng.set_fused(fused)
        tuned = dict(tuned or {})
        tuned["fused_sweep"] = a.fused
    elif a.fused == "auto":
        from hormone_amd.batch import tune_fused
        fused, ftimes = tune_fused(eng, run.advance, steps=max(2, min(8, a.warmup // 4)))
        tuned = dict(tuned or {})
        tuned["fused_sweep"] = names[fused]
        tuned["ms_per_step"] = {names[k]: round(v * 1e3, 3) for k, v in ftimes.items()}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ok = run.advance(a.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert ok, "the run ended inside the timed region; lower --steps"
    zus = a.steps * g.ncells / wall
    roof = None
    if not a.no_profile:
        # The timed region carries no per-dispatch events: on this stack each
        # one adds a completion wait of about 9 us to its launch (15% of an
        # OT 512^2 step).  The sweep's own duration is timed right after, as
        # back-to-back launches on the same state and stream between two HIP
        # events (the sweep only rewrites the flux divergences); the other
        # kernels over a short window with per-dispatch events.
        _, _, bpl = eng.profile_read(native.PROF_SWEEP)
        est = eng.time_sweep(2)
        nrep = int(min(2000, max(20, 0.25e3 / max(est / 2, 1e-3))))
        ms = eng.time_sweep(nrep)
        avg_s = ms / 1e3 / nrep
        ach = bpl / avg_s / 1e9
        eng.profile(True)
        win = min(20, a.steps)
        run.advance(win)
        ev_ms, ev_n, _ = eng.profile_read(native.PROF_SWEEP)
        rk_ms, rk_n, rk_b = eng.profile_read(native.PROF_RK)
        gr_ms, gr_n, gr_b = eng.profile_read(native.PROF_GRAV)
        eng.profile(False)
        kname = "k_sweep" + ("_mix" if a.precision == "fp32mix" else "")
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(a.workload, a.precision, kname),
                "kernel": kname,
                "kernel_note": "all solved directions in one launch per RK stage; limited by "
                               "FP64 VALU latency (IEEE division/square-root chains), not HBM",
                "avg_launch_us": round(avg_s * 1e6, 2), "launches": nrep,
                "timing": "back-to-back launches between two HIP events, after the timed region",
                "algorithmic_bytes_per_launch": bpl,
                "share_of_timed_region": round(avg_s * a.steps * prm.rktype / wall, 3)}
        if fused:
            # the timed region ran one direction's sweep fused with the RK
            # update (k_sweep_rk) and the other directions in k_sweep; the
            # replay above times the all-direction k_sweep
            roof["timed_region_form"] = {1: "last", 2: "x"}[fused] + " direction fused with RK"
            roof.pop("share_of_timed_region")
        if ev_n:  # fused: the sweep of the other directions only
            roof["other_directions_sweep_avg_launch_us" if fused else "event_avg_launch_us"] = \
                round(ev_ms / ev_n * 1e3, 2)
        if rk_n and fused:  # the RK events time the fused sweep + RK launch
            roof["fused_sweep_rk_avg_launch_us"] = round(rk_ms / rk_n * 1e3, 2)
        elif rk_n:
            roof["rk_avg_launch_us"] = round(rk_ms / rk_n * 1e3, 2)
            roof["rk_achieved_GBs"] = round(rk_b / (rk_ms / rk_n / 1e3) / 1e9, 1)
        if gr_n:
            roof["grav_avg_launch_us"] = round(gr_ms / gr_n * 1e3, 2)
            roof["grav_achieved_GBs"] = round(gr_b / (gr_ms / gr_n / 1e3) / 1e9, 1)
    parity = None
    extra = {}
    # whole-step figure: compulsory FP64 state traffic per zone-update (RK3
    # reads/writes the conserved state: 72 B x nvar; with hyperbolic gravity
    # + 24 B potential reads + 8 B source + 96 B per gravity sub-step)
    b_zu = 72.0 * eng.ufnmax()
    if run.nsub:
        b_zu += 24.0 + 8.0 + 96.0 * float(np.mean(run.nsub[-a.steps:]))
    extra["step_roofline"] = {"bytes_per_zone_update": b_zu,
                              "achieved": round(zus * b_zu / 1e9, 1), "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": round(zus * b_zu / 1e9 / HBM_PEAK_GBS, 4)}
    if a.workload == "ot512" and a.precision == "fp64" and not a.no_check and (a.n or 512) == 512:
        from tests.helpers import meta
        m = meta()
        run.advance(2200 - run.tn)
        out = D.fetch(eng, g, D.SNAP_FIELDS_MAG)
        e = m.get("ot512_2200steps")
        same = e is not None and run.tn == e["tn"] and run.tim == e["time"] and all(
            hashlib.sha256(np.ascontiguousarray(out[f], "<f8").tobytes()).hexdigest()
            == e["fields"][nm]["sha256"] for f, nm in zip(D.SNAP_FIELDS_MAG, NAMES))
        parity = {"step": run.tn, "bitwise_equal_to_reference_run_at_step_2200": bool(same)}
        fail = m.get("ot512_t05_failure")
        try:
            run.advance(1 << 30)
            parity["reference_stop_reproduced"] = False
        except native.HgError as err:
            parity["stop"] = {"code": err.code, "completed_steps": err.tn, "cell": list(err.cell)}
            parity["reference_stop_reproduced"] = bool(
                fail and err.code == native.HG_ERR_NEG_EINT
                and err.tn == fail["completed_steps"] and list(err.cell) == fail["cell"])
    if a.workload != "ot512" or a.precision != "fp64":
        m1, e1 = sums(D, eng, g, prm)
        extra["mass_drift"] = (m1 - m0) / m0
        extra["energy_drift"] = (e1 - e0) / abs(e0)
        if run.nsub:
            extra["gravity_substeps_mean"] = float(np.mean(run.nsub))
    if a.vs_fp64 and a.precision == "fp32mix":
        fields = D.SNAP_FIELDS_MAG if prm.mag_on else D.SNAP_FIELDS
        mine = D.fetch(eng, g, fields)
        n_done = run.tn
        del eng
        prm64, g64, st64 = setup(a, "fp64")
        eng64 = native.Hydro(**D.desc_fields(prm64, g64))
        D.load_initial_state(eng64, prm64, g64, st64)
        # compare at the same physical time: the mixed run's dt sequence
        # differs slightly (its CFL speeds carry FP32 rounding), so the FP64
        # run is stepped to the mixed run's clock (last dt clipped onto it)
        r64 = Runner(eng64, prm64, g64)
        advance_to(r64, run.tim)
        ref = D.fetch(eng64, g64, fields + (D.F_P,))
        extra["rel_l1_vs_fp64"] = D.rel_l1_errors(mine, ref)
        extra["reference_norms_vs_fp64"] = D.reference_norms(
            mine, {f: ref[f] for f in fields}, g64, prm.gamma, bool(prm.mag_on))
        # velocity error in units of the local sound speed (the relevant
        # scale for a near-hydrostatic flow, where |v| itself is noise)
        cs = np.sqrt(prm.gamma * ref[D.F_P] / ref[D.F_D])
        extra["vel_l1_over_cs"] = float(sum(np.abs(mine[f] - ref[f]).sum() for f in
                                            (D.F_V1, D.F_V2, D.F_V3)) / (3.0 * cs.sum()))
        extra["rel_l1_steps"] = n_done
        extra["rel_l1_time"] = run.tim
        extra["fp64_steps_to_same_time"] = r64.tn
    cpu = None if a.no_cpu else cpu_baseline(a, a.cpu_seconds)
    wl = {"ot512": f"orszagtang_xy {g.ax[0].n}x{g.ax[1].n} MHD+GLM, MUSCL+HLLD, RK3, "
                   f"courant 0.9 (run toward t=0.5)",
          "polytrope256": f"polytrope {g.ax[0].n}^3 {'MHD B=0' if prm.mag_on else 'hydro'} + "
                          f"hyperbolic self-gravity, gamma 2, {prm.tnlim} steps",
          "kh384": f"Kelvin-Helmholtz {g.ax[0].n}^3 MHD, periodic, {prm.tnlim} steps"}
    line = {
        "metric": f"zone-updates/s ({a.precision.upper()}, {a.workload})",
        "value": zus, "unit": "zone-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic (closed-form initial state)",
        "config": {"workload": wl[a.workload], "cells": g.ncells, "precision": a.precision},
        "roofline": roof, "cpu_baseline": cpu, "parity": parity,
        "launch_tuning": tuned,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
