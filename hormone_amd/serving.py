"""Continuous-batching serving loop for independent simulations.

Requests (2D/3D Riemann or blast problems of different grid sizes) arrive
over time; the loop keeps their state resident on the device, admits a
queued request as soon as it has arrived and its memory fits the HBM budget,
advances the admitted requests of one size class together through one
:class:`Batch` (their work overlaps on the device in the same launches; the
classes run concurrently on their own streams), retires each one when it
reaches its end time and frees its memory for the next.  It reports p50/p99
time-to-solution (arrival to completion) and aggregate zone-updates/s.

Each request's result is bit-identical alone and in the mix: boxes share
launches, never data, and every box keeps its own dt.
"""
from __future__ import annotations

import hashlib
import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import driver as D
from . import problems as P
from .batch import Batch, autotune, shape_class
from .grid import cartesian_grid
from .native import Hydro, HgError

GAMMA = 5.0 / 3.0


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the run
    kind: str               # "riemann" or "blast"
    dim: int
    n: int
    rho: tuple
    p: tuple
    v: tuple                # uniform velocity
    b_left: tuple
    b_right: tuple
    t_end: float = 0.1
    # filled by the server
    admitted: float | None = None
    finished: float | None = None
    steps: int = 0
    error: str | None = None
    digest: str | None = None

    @property
    def cells(self) -> int:
        return self.n ** self.dim

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


def make_requests(count: int, rate: float = 50.0, seed: int = 7,
                  min_cells: int = 64 ** 2, max_cells: int = 192 ** 3) -> list:
    """Seeded Poisson arrivals; grid sizes log-uniform in cells from 64^2 to
    192^3 (2D up to 192^2, 3D above); rho and p lognormal(0,1); a uniform
    velocity with |v| <= 1; B components uniform in [-1,1] (normal component
    continuous across the Riemann interface); gamma 5/3; t_end = 0.1."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    out = []
    for q in range(count):
        cells = 10.0 ** rng.uniform(math.log10(min_cells), math.log10(max_cells))
        if cells <= 192 ** 2:
            dim, n = 2, int(round(math.sqrt(cells)))
        else:
            dim, n = 3, int(round(cells ** (1.0 / 3.0)))
        n = max(16, n)
        kind = "riemann" if rng.uniform() < 0.5 else "blast"
        rho = tuple(rng.lognormal(0.0, 1.0, 2))
        p = tuple(rng.lognormal(0.0, 1.0, 2))
        direc = rng.normal(size=3)
        direc /= np.linalg.norm(direc)
        v = tuple(direc * rng.uniform(0.0, 1.0))
        bl = rng.uniform(-1.0, 1.0, 3)
        br = rng.uniform(-1.0, 1.0, 3)
        br[0] = bl[0]  # B_x continuous across the x-interface
        if kind == "blast":
            br = bl.copy()
        out.append(Request(q, float(t[q]), kind, dim, n, rho, p, v, tuple(bl), tuple(br)))
    return out


def setup(req: Request):
    """(Params, Grid, State) of one request: [0,1]^dim, free boundaries,
    MUSCL + HLLD/GLM MHD, courant 0.9."""
    nn = (req.n, req.n, req.n if req.dim == 3 else 1)
    prm = P.Params(n=nn, gamma=GAMMA, mag_on=True, bc=(3,) * 12, t_end=req.t_end,
                   dt_out=req.t_end, dt_unit="s", is_test=True)
    g = cartesian_grid(nn, prm.lo, prm.hi)
    x = g.centres(0)[None, None, :]
    y = g.centres(1)[None, :, None]
    z = g.centres(2)[:, None, None]
    if req.kind == "riemann":
        left = np.broadcast_to(x <= 0.5, g.shape)
        rho = np.where(left, req.rho[0], req.rho[1])
        pr = np.where(left, req.p[0], req.p[1])
        b = [np.where(left, req.b_left[a], req.b_right[a]) for a in range(3)]
    else:
        r2 = (x - 0.5) ** 2 + (y - 0.5) ** 2 + ((z - 0.5) ** 2 if req.dim == 3 else 0.0)
        inside = np.broadcast_to(r2 < 0.1 ** 2, g.shape)
        rho = np.full(g.shape, req.rho[0])
        pr = np.where(inside, 10.0 * max(req.p), req.p[1])
        b = [np.full(g.shape, req.b_left[a]) for a in range(3)]
    v = [np.full(g.shape, req.v[a]) for a in range(3)]
    st = P.State(d=np.ascontiguousarray(rho, dtype=float), p=np.ascontiguousarray(pr, dtype=float),
                 v=v, b=[np.ascontiguousarray(q, dtype=float) for q in b],
                 phi=np.zeros(g.shape))
    return prm, g, st


def device_bytes(req: Request, batch_fused: bool = True) -> int:
    """HBM footprint of one MHD context: the 5 W fields that do not alias U,
    U 9, U0 9, DU dim*9, and in an x-fused batch the second state set (U 9,
    W 5) and the third set's own W fields (5; its U is U0's arrays)."""
    nv = 9
    return 8 * req.cells * (5 + 2 * nv + req.dim * nv + (nv + 10 if batch_fused else 0)) + \
        (1 << 20)


def state_digest(eng, g) -> str:
    out = D.fetch(eng, g, D.SNAP_FIELDS_MAG)
    h = hashlib.sha256()
    for f in D.SNAP_FIELDS_MAG:
        h.update(np.ascontiguousarray(out[f], "<f8").tobytes())
    return h.hexdigest()


def run_alone(req: Request) -> tuple:
    """The request on its own through hg_run: (steps, time, digest)."""
    prm, g, st = setup(req)
    eng = Hydro(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    r = D.run(eng, prm, g, device_loop=True)
    return r.tn, r.time, state_digest(eng, g)


@dataclass
class ServeReport:
    requests: list
    makespan: float
    zone_updates: int
    tuned_shape: tuple | None = None
    events: list = field(default_factory=list)

    def latencies(self):
        return np.array([r.latency for r in self.requests if r.latency is not None])

    def summary(self) -> dict:
        lat = self.latencies()
        done = [r for r in self.requests if r.finished is not None and r.error is None]
        return {
            "requests": len(self.requests), "completed": len(done),
            "failed": sum(r.error is not None for r in self.requests),
            "p50_s": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_s": float(np.percentile(lat, 99)) if lat.size else None,
            "aggregate_zone_updates_per_s": self.zone_updates / self.makespan,
            "makespan_s": self.makespan, "tuned_shape": self.tuned_shape,
        }


class _Shared:
    """State the size-class lanes share: the HBM budget, the zone-update
    count and the event log."""

    def __init__(self, budget: float):
        self.lock = threading.Lock()
        self.budget = budget
        self.used = 0
        self.zu = 0
        self.events = []
        self.tuned = []


def _serve_lane(lane: int, queue: list, sh: _Shared, t0: float, steps_per_round: int,
                tune: bool, keep_digest: bool, digest_ids, batch_fused: bool = True):
    """One size class: its own Batch (hence its own stream and hipGraph),
    admitting its requests in arrival order as the shared budget allows."""
    qi = 0
    batch = Batch()
    batch.set_fused(batch_fused)
    active = {}   # id(eng) -> (req, eng, g)
    tuned_class = None
    try:
        while qi < len(queue) or active:
            now = time.perf_counter() - t0
            while qi < len(queue) and queue[qi].arrival <= now:
                req = queue[qi]
                need = device_bytes(req, batch_fused)
                with sh.lock:
                    # over the budget, wait for memory unless nothing at all
                    # is in flight (a request larger than the budget still runs)
                    if sh.used + need > sh.budget and sh.used > 0:
                        break
                    sh.used += need
                prm, g, st = setup(req)
                eng = Hydro(**D.desc_fields(prm, g))
                D.load_initial_state(eng, prm, g, st)
                t_end, dt_out, t_out = prm.times()
                batch.add(eng, t_end=t_end, t_out=t_out, dt_out=dt_out)
                req.admitted = time.perf_counter() - t0
                active[id(eng)] = (req, eng, g)
                qi += 1
            if not active:
                if qi < len(queue):
                    time.sleep(max(0.0, min(0.005, queue[qi].arrival - now)))
                continue
            # (re)tune the launch shapes whenever the mix has moved to another
            # size class (tuned shapes are cached per class)
            if tune and len(active) >= 2:
                cls = shape_class(batch)
                if cls != tuned_class:
                    tuned, _ = autotune(batch)
                    tuned_class = cls
                    with sh.lock:
                        sh.tuned.append((lane, tuned))
            try:
                batch.step(steps_per_round)
            except HgError as err:
                # the batch step itself failed (a launch, a graph rebuild or
                # its memory): every request in flight fails with it and
                # leaves the batch, so the lane never spins on a broken batch
                for key, (req, eng, g) in list(active.items()):
                    req.error = f"batch step: {err}"
                    req.finished = time.perf_counter() - t0
                    batch.remove(eng)
                    eng.close()
                    with sh.lock:
                        sh.used -= device_bytes(req, batch_fused)
                    del active[key]
                continue
            with sh.lock:
                sh.events.append({"t": round(time.perf_counter() - t0, 4), "lane": lane,
                                  "active": len(active),
                                  "queued": sum(1 for r in queue[qi:] if r.arrival <= now),
                                  "hbm_bytes": sh.used})
            for key, (req, eng, g) in list(active.items()):
                try:
                    tm, tn, stop = batch.state(eng)
                except HgError as err:
                    req.error = str(err)
                    tm, tn, stop = None, 0, 4
                if stop != 0:
                    req.finished = time.perf_counter() - t0
                    req.steps = tn
                    batch.remove(eng)
                    if stop == 1 and req.error is None:
                        with sh.lock:
                            sh.zu += req.cells * tn
                        if keep_digest or (digest_ids is not None and req.rid in digest_ids):
                            req.digest = state_digest(eng, g)
                    eng.close()
                    with sh.lock:
                        sh.used -= device_bytes(req, batch_fused)
                    del active[key]
    finally:
        batch.close()


def serve(requests: list, hbm_budget: float = 64 * 2 ** 30, steps_per_round: int = 4,
          tune: bool = True, keep_digest: bool = False, digest_ids=None,
          lanes=(50_000, 500_000), batch_fused: bool = True) -> ServeReport:
    """Continuous batching over arriving requests (arrival times are honoured
    in wall-clock time from the start of the call).  A final-state digest is
    kept for every request (keep_digest) or for the ids in digest_ids.

    lanes: upper cell-count bounds of the size classes below the largest.
    Each class has its own batch -- own stream, hipGraph and host thread --
    so small requests advance at their own step rate instead of one step per
    step of the largest box in flight, while all classes share the GPU and
    the HBM budget.  lanes=() serves everything from one batch.
    batch_fused: the batches' x-fused form (hg_batch_set_fused) or the
    separate RK update; the same bits either way."""
    bounds = sorted(lanes) + [math.inf]
    queues = [[] for _ in bounds]
    for r in sorted(requests, key=lambda r: r.arrival):
        queues[next(k for k, b in enumerate(bounds) if r.cells <= b)].append(r)
    sh = _Shared(hbm_budget)
    errors = []

    def worker(lane, queue):
        try:
            _serve_lane(lane, queue, sh, t0, steps_per_round, tune, keep_digest, digest_ids,
                        batch_fused)
        except BaseException as err:  # re-raised below
            errors.append(err)

    t0 = time.perf_counter()
    threads = [threading.Thread(target=worker, args=(k, q)) for k, q in enumerate(queues) if q]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    makespan = time.perf_counter() - t0
    if errors:
        raise errors[0]
    return ServeReport(requests, makespan, sh.zu, tuple(sh.tuned), sh.events)
