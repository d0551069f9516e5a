"""Batches of independent boxes and the launch-shape autotuner.

A :class:`Batch` owns an ``hg_batch``: many device-resident contexts
(:class:`hormone_amd.native.Hydro`) of uneven sizes that advance together,
one persistent launch per kernel per RK stage over flattened work-lists,
replayed from a captured hipGraph.  Each box keeps its own dt and clock, so
its result is bit-identical to running it alone.

:func:`autotune` picks the launch shapes (sweep wavefronts per block and the
persistent grid sizes of the sweep and RK launches) by timing real steps of
the batch itself.  The shapes change only which wavefront computes which
cell, never the arithmetic, so tuning steps are ordinary steps of the run.
Results are cached per shape class (rounded total cells and box count).
"""
from __future__ import annotations

import ctypes as C
import math
import time

from . import native

_dp = C.POINTER(C.c_double)
_bound = False


def _bind():
    global _bound
    if _bound:
        return native.lib()
    L = native.lib()
    vp = C.c_void_p
    L.hg_batch_create.argtypes = [C.c_int32, C.POINTER(vp)]
    L.hg_batch_destroy.argtypes = [vp]
    L.hg_batch_add.argtypes = [vp, vp, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.c_int32, C.c_int32, C.c_int64, C.c_int64]
    L.hg_batch_remove.argtypes = [vp, vp]
    L.hg_batch_step.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
    L.hg_batch_member_state.argtypes = [vp, vp, _dp, C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int32)]
    L.hg_batch_config.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    L.hg_batch_set_fused.argtypes = [vp, C.c_int32]
    L.hg_batch_form.argtypes = [vp, C.POINTER(C.c_int32)]
    _bound = True
    return L


class Batch:
    def __init__(self, device: int = 0):
        self._lib = _bind()
        h = C.c_void_p()
        st = self._lib.hg_batch_create(device, C.byref(h))
        if st:
            raise native.HgError(st, "hg_batch_create")
        self.h = h
        self.members = []
        self.config = (4, 0, 0, 1)

    def close(self):
        if getattr(self, "h", None):
            self._lib.hg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, eng: native.Hydro, t_end: float, t_out: float | None = None,
            dt_out: float = 1.0, outstyle: int = 1, endstyle: int = 1, tnlim: int = 0,
            tn_out: int = 1, time0: float = 0.0):
        """Add a context; by default it runs to t_end with t_out = t_end."""
        if t_out is None:
            t_out = t_end
        st = self._lib.hg_batch_add(self.h, eng.h, time0, t_out, dt_out, t_end, outstyle,
                                    endstyle, tnlim, tn_out)
        if st:
            raise native.HgError(st, "hg_batch_add")
        self.members.append(eng)

    def remove(self, eng: native.Hydro):
        st = self._lib.hg_batch_remove(self.h, eng.h)
        if st:
            raise native.HgError(st, "hg_batch_remove")
        self.members.remove(eng)

    def step(self, nsteps: int = 1) -> int:
        n = C.c_int32()
        st = self._lib.hg_batch_step(self.h, nsteps, C.byref(n))
        if st:
            raise native.HgError(st, "hg_batch_step")
        return n.value

    def state(self, eng: native.Hydro):
        """(time, tn, stop) of a member after the last step."""
        t, n, s = C.c_double(), C.c_int64(), C.c_int32()
        st = self._lib.hg_batch_member_state(self.h, eng.h, C.byref(t), C.byref(n), C.byref(s))
        if st:
            raise native.HgError(st, "hg_batch_member_state")
        return t.value, n.value, s.value

    def set_fused(self, on: bool = True):
        """x sweep with the RK update (the default, when every member sweeps
        x) or the separate RK update; the same bits either way."""
        st = self._lib.hg_batch_set_fused(self.h, int(bool(on)))
        if st:
            raise native.HgError(st, "hg_batch_set_fused")

    def fused(self) -> bool:
        """Whether the launches were last built in the x-fused form."""
        f = C.c_int32()
        st = self._lib.hg_batch_form(self.h, C.byref(f))
        if st:
            raise native.HgError(st, "hg_batch_form")
        return bool(f.value)

    def configure(self, sweep_block_waves=4, sweep_grid=0, rk_grid=0, use_graph=True):
        st = self._lib.hg_batch_config(self.h, sweep_block_waves, sweep_grid, rk_grid,
                                       int(use_graph))
        if st:
            raise native.HgError(st, "hg_batch_config")
        self.config = (sweep_block_waves, sweep_grid, rk_grid, int(use_graph))


# --- autotuner ------------------------------------------------------------------

CUS = 256  # MI355X compute units
_CACHE: dict = {}


def shape_class(batch: Batch):
    """Coarse class of a batch for the tuning cache: member count and total
    cells, each rounded to a power of two."""
    import math
    cells = sum(m.grid.ncells for m in batch.members)
    n = len(batch.members)
    return (round(math.log2(max(n, 1))), round(math.log2(max(cells, 1))))


def _time_steps(batch: Batch, cand, steps: int):
    batch.configure(*cand, use_graph=True)
    batch.step(1)  # capture the graph outside the timing
    t0 = time.perf_counter()
    if batch.step(steps) == 0:
        return None
    return (time.perf_counter() - t0) / steps


def autotune(batch: Batch, steps_per_trial: int = 2, use_cache: bool = True):
    """Pick the launch shape of a batch by timing real steps of it, in three
    stages: sweep wavefronts per block (4, 8, 16) with one work item per
    wavefront; then the persistent sweep grid; then the RK grid (the
    separate form only).  Every
    trial step is an ordinary step of the run (the shape never changes the
    arithmetic).  Returns (best (block waves, sweep grid, rk grid),
    {shape: seconds per step})."""
    key = shape_class(batch)
    if use_cache and key in _CACHE:
        best = _CACHE[key]
        batch.configure(*best, use_graph=True)
        return best, {}
    timings = {}

    def trial(c):
        if c not in timings:
            t = _time_steps(batch, c, steps_per_trial)
            if t is None:
                return False
            timings[c] = t
        return True

    for bw in (4, 8, 16):
        if not trial((bw, 0, 0)):
            break
    if timings:
        bw = min(timings, key=timings.get)[0]
        for sg in (CUS * 4, CUS * 8, CUS * 16):
            if not trial((bw, sg, 0)):
                break
        best = min(timings, key=timings.get)
        # the x-fused form has no separate RK launch
        for rg in (() if batch.fused() else (CUS * 4, CUS * 8)):
            if not trial((best[0], best[1], rg)):
                break
    best = min(timings, key=timings.get) if timings else (4, 0, 0)
    batch.configure(*best, use_graph=True)
    _CACHE[key] = best
    return best, timings


def _time_host_steps(advance, steps: int, min_s: float = 0.02, cap: int = 20) -> float:
    """Seconds per step over at least `steps` steps after one untimed step,
    more (up to `cap`) when that would take under `min_s`: host-clock noise
    dominates short measurements of small boxes."""
    t0 = time.perf_counter()
    advance(1)
    est = time.perf_counter() - t0
    n = max(steps, min(cap, int(math.ceil(min_s / max(est, 1e-6)))))
    t0 = time.perf_counter()
    advance(n)
    return (time.perf_counter() - t0) / n


def tune_block_steps(eng: native.Hydro, advance, candidates=(2, 4, 8, 16), steps: int = 2):
    """Pick the sweep block size by timing whole steps (host clock) in the
    form the context runs now, so a fused sweep + RK launch counts too;
    ``advance(k)`` must take k steps and return after they finish.  Returns
    (best, {candidate: s/step})."""
    timings = {}
    for bw in candidates:
        eng.set_launch(bw)
        timings[bw] = _time_host_steps(advance, steps)
    best = min(timings, key=timings.get)
    eng.set_launch(best)
    return best, timings


def tune_fused(eng: native.Hydro, advance, steps: int = 4):
    """Choose between the separate RK update and a sweep direction fused
    with it (hg_set_fused: 1 the last solved direction, 2 x) by timing
    whole steps of each, on the host clock; ``advance(k)`` must take k steps
    and return after they finish.  Returns (mode, {mode: s/step}) with mode
    0 for the separate update; the modes the context does not support are
    skipped.  Every form gives the same bits, so the tuning steps are
    ordinary steps of the run."""
    modes = [0]
    for m in (1, 2):
        try:
            eng.set_fused(m)
            modes.append(m)
        except native.HgError as err:
            if err.code != native.HG_ERR_UNSUPPORTED:
                raise
    eng.set_fused(0)
    if len(modes) == 1:
        return 0, {}
    timings = {}
    for _ in range(2):
        for m in modes:
            eng.set_fused(m)
            dt = _time_host_steps(advance, steps)
            timings[m] = min(dt, timings.get(m, dt))
    best = min(timings, key=timings.get)
    eng.set_fused(best)
    return best, timings
