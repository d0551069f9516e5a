"""Host-side main loop over a hydro engine, mirroring the reference driver.

``run`` restates the part of ``program hormone`` (src/main/hormone.f90:61-179)
that touches the hot path: initial state (initialcondition + conserve +
pressure), then per step ``timestep`` -> [``gravity``] -> ``hydro_step`` ->
``phidamp`` (when mag_on and dim>1) -> ``time += dt`` with the reference's
end and output bookkeeping (endstyle/outstyle, t_out += dt_out).

An engine is any object with the methods of :class:`hormone_amd.native.Hydro`
(upload, download, timestep, hydro_step, phidamp, ufnmax); the HIP library
is the product engine.
"""
from __future__ import annotations

import math
import time as _time
from dataclasses import dataclass, field

import numpy as np

from . import problems as P
from .grid import Grid

F_D, F_V1, F_V2, F_V3, F_E, F_B1, F_B2, F_B3, F_PHI, F_EINT, F_P = range(11)
F_GRVPHI, F_GRVPSI, F_U, F_SPC = 11, 12, 16, 32


def desc_fields(prm: P.Params, g: Grid) -> dict:
    """Keyword values for the C descriptor (hg_desc)."""
    return dict(grid=g, mag_on=int(prm.mag_on),
                flux_limiter=P.LIMITERS[prm.flux_limiter], rktype=prm.rktype,
                bc=prm.effective_bc(), gamma=prm.gamma, courant=prm.courant,
                alpha9wave=prm.alpha9wave, gravswitch=prm.gravswitch,
                grktype=prm.grktype, HGfac=prm.HGfac, hgcfl=prm.hgcfl,
                alphagrv=prm.alphagrv, G=prm.G, maxtngrv=prm.maxtngrv,
                precision=1 if prm.precision == "fp32mix" else 0,
                compswitch=prm.compswitch, spn=prm.spn)


def load_initial_state(eng, prm: P.Params, g: Grid, st: P.State):
    """initialcondition (energy) + conserve; the engine recomputes eint/p
    from e, d, v, B exactly as pressure() does before the first step."""
    eint, e = P.initial_energy(st, prm.gamma)
    u = P.conserve(st, e, prm.mag_on, g.dim)
    prims = {F_D: st.d, F_V1: st.v[0], F_V2: st.v[1], F_V3: st.v[2], F_E: e,
             F_B1: st.b[0], F_B2: st.b[1], F_B3: st.b[2], F_PHI: st.phi}
    for f, a in prims.items():
        eng.upload(f, g.pad(a))
    for n, a in enumerate(u):
        eng.upload(F_U + n, g.pad(a))
    if prm.compswitch == 2:
        for n, a in enumerate(st.extra["spc"]):
            eng.upload(F_SPC + n, g.pad(a))
    if prm.gravswitch == 3:
        eng.upload(F_GRVPHI, st.extra["grvphi_padded"])
        eng.upload(F_GRVPSI, np.zeros(g.padded_shape))


def load_restart_state(eng, prm: P.Params, g: Grid, snap: dict, grav_boundary=None):
    """Restart from a snapshot (restart.f90:15-64 -> readbin.f90:16-115,
    then hormone.f90:86-90): the primitives and total energy as read, the
    conserved variables from them (conserve), e_int from them (pressure, by
    the engine).  Returns (tn, time) of the snapshot, and cgrav_old as a
    third element for hyperbolic gravity (gravswitch 3), whose potential
    and psi come from the snapshot too; grav_boundary is then the padded
    potential the run started with, whose ghost layer holds the boundary
    values."""
    sh = g.shape
    d = np.asarray(snap["d"]).reshape(sh)
    v = [np.asarray(snap[k]).reshape(sh) for k in ("v1", "v2", "v3")]
    e = np.asarray(snap["e"]).reshape(sh)
    z = np.zeros(sh)
    b = [np.asarray(snap[k]).reshape(sh) for k in ("b1", "b2", "b3")] if prm.mag_on else [z, z, z]
    phi = np.asarray(snap["phi"]).reshape(sh) if prm.mag_on else z
    st = P.State(d=d, p=z, v=v, b=b, phi=phi)
    prims = {F_D: d, F_V1: v[0], F_V2: v[1], F_V3: v[2], F_E: e,
             F_B1: b[0], F_B2: b[1], F_B3: b[2], F_PHI: phi}
    for f, a in prims.items():
        eng.upload(f, g.pad(a))
    for n, a in enumerate(P.conserve(st, e, prm.mag_on, g.dim)):
        eng.upload(F_U + n, g.pad(a))
    if prm.compswitch == 2:
        for q in range(prm.spn):
            eng.upload(F_SPC + q, g.pad(np.asarray(snap[f"spc{q + 1}"]).reshape(sh)))
    if prm.gravswitch == 3:
        phi = np.array(grav_boundary, dtype=np.float64, copy=True)
        phi[2:-2, 2:-2, 2:-2] = np.asarray(snap["grvphi"]).reshape(sh)
        eng.upload(F_GRVPHI, phi)
        eng.upload(F_GRVPSI, g.pad(np.asarray(snap["grvpsi"]).reshape(sh)))
        return int(snap["tn"]), float(snap["time"]), float(snap["cgrav_old"])
    return int(snap["tn"]), float(snap["time"])


def rel_l1_errors(mine: dict, ref: dict) -> dict:
    """Relative L1 error of each snapshot field against a reference run.

    Scalars (d, e, phi): sum|q - q_ref| / sum|q_ref|.  Vector components are
    normalised by the reference's vector magnitude, as the reference's
    regression norm does (tests.f90:141-153: |v| for v1..v3, |B| for
    b1..b3), so a component that happens to be near zero everywhere does not
    read as a large relative error: sum|v_i - v_i,ref| / sum|v_ref|.
    """
    def mag(fs):
        return np.sqrt(sum(ref[f] * ref[f] for f in fs))

    scale = {}
    vmag = mag((F_V1, F_V2, F_V3))
    for f in (F_V1, F_V2, F_V3):
        scale[f] = vmag
    if F_B1 in ref:
        bmag = mag((F_B1, F_B2, F_B3))
        for f in (F_B1, F_B2, F_B3):
            scale[f] = bmag
    names = {F_D: "d", F_V1: "v1", F_V2: "v2", F_V3: "v3", F_E: "e", F_B1: "b1",
             F_B2: "b2", F_B3: "b3", F_PHI: "phi"}
    out = {}
    for f, a in mine.items():
        den = float(np.abs(scale.get(f, ref[f])).sum())
        num = float(np.abs(a - ref[f]).sum())
        out[names.get(f, str(f))] = num / den if den > 0 else num
    return out


def reference_norms(mine: dict, ref: dict, g: Grid, gamma: float, mag_on: bool) -> dict:
    """The reference's own test norms (tests.f90:198-298: max, "L1", and
    discontinuity-weighted L2) of a snapshot against another run, keyed
    {norm: {variable: error}}; see hormone_amd/testnorms.py."""
    from . import testnorms as T
    names = {F_D: "d", F_V1: "v1", F_V2: "v2", F_V3: "v3", F_E: "e", F_B1: "b1",
             F_B2: "b2", F_B3: "b3", F_PHI: "phi"}
    sh = g.shape
    a = {names[f]: np.asarray(v).reshape(sh) for f, v in mine.items() if f in names}
    b = {names[f]: np.asarray(v).reshape(sh) for f, v in ref.items() if f in names}
    _, errs, _ = T.reference_criterion(a, b, gamma, T.cell_volumes(g), 0.0, mag_on, g.dim)
    return errs


@dataclass
class RunResult:
    tn: int
    time: float
    wall_s: float
    steps_timed: int
    dts: list = field(default_factory=list)
    nsub: list = field(default_factory=list)  # gravity sub-steps per step
    cgrav_old: float = 0.0  # gravmod's cgrav_old at the end (gravswitch 3)


def run(eng, prm: P.Params, g: Grid, max_steps: int | None = None,
        record_dt: bool = False, device_loop: bool = False, start=(0, 0.0)) -> RunResult:
    """Main loop of hormone.f90:124-179 restricted to the hydro path.

    ``device_loop`` runs it on the device (hg_run) when the engine offers it;
    the state is the same bit for bit.  ``start`` = (tn, time) of a restart
    (load_restart_state); max_steps counts from step 0 like tnlim.
    """
    t_end, dt_out, t_out = prm.times()
    if (device_loop and hasattr(eng, "run_loop") and prm.fixed_dt <= 0 and not record_dt
            and prm.gravswitch == 0):
        return _run_device(eng, prm, g, max_steps, t_end, dt_out, t_out, start)
    if (device_loop and hasattr(eng, "run_loop_gravity") and prm.fixed_dt <= 0
            and not record_dt and prm.gravswitch == 3):
        return _run_device_gravity(eng, prm, g, max_steps, t_end, dt_out, t_out, start)
    tn, tim = int(start[0]), float(start[1])
    damp = math.exp(-prm.alpha9wave * prm.courant)  # source.f90:253
    dts = []
    nsubs = []
    grav = prm.gravswitch == 3
    grvtime, cgrav_old = 0.0, 0.0  # gravmod module values at start-up
    if grav and len(start) > 2:
        # restart: cgrav_old from the snapshot, grvtime = time (restart.f90:59)
        grvtime, cgrav_old = tim, float(start[2])
    t0 = _time.perf_counter()
    while True:
        if grav:
            dt, ch, cgrav, dtgrav = eng.timestep(tim, t_out, t_end, prm.outstyle,
                                                 prm.endstyle, prm.fixed_dt, grav=True)
            # gravity.f90:41-60: at tn == 0 the potential is the initial one
            # (grav_init_other); afterwards the hyperbolic solver advances it.
            if tn == 0:
                grvtime = 0.0
            else:
                grvtime, cgrav_old, dtgrav, ns = eng.gravity(tim, dt, grvtime, cgrav,
                                                             cgrav_old, dtgrav)
                nsubs.append(ns)
        else:
            dt, ch = eng.timestep(tim, t_out, t_end, prm.outstyle, prm.endstyle,
                                  prm.fixed_dt)
        if record_dt:
            dts.append(dt)
        eng.hydro_step(dt, ch)
        if prm.mag_on and g.dim > 1:
            eng.phidamp(damp)
        tim = tim + dt
        tn += 1
        if prm.endstyle == 1:
            if tim >= t_end:
                break
        elif prm.endstyle == 2:
            if tn >= prm.tnlim:
                break
        if prm.outstyle == 1:
            if tim >= t_out:
                t_out = t_out + dt_out
        if max_steps is not None and tn >= max_steps:
            break
    wall = _time.perf_counter() - t0
    return RunResult(tn, tim, wall, tn, dts, nsubs, cgrav_old)


def _run_device(eng, prm, g, max_steps, t_end, dt_out, t_out, start=(0, 0.0)) -> RunResult:
    damp = math.exp(-prm.alpha9wave * prm.courant) if (prm.mag_on and g.dim > 1) else 0.0
    tn, tim = int(start[0]), float(start[1])
    budget = max_steps if max_steps is not None else 1 << 62
    t0 = _time.perf_counter()
    while True:
        tim, tn, t_out, reason = eng.run_loop(tim, tn, t_out, dt_out, t_end, prm.outstyle,
                                              prm.endstyle, prm.tnlim, prm.tn_out, budget - tn,
                                              damp)
        # reason 2: an output is due (none is written here); 1: end; 3: max steps
        if reason != 2 or tn >= budget:
            break
    wall = _time.perf_counter() - t0
    return RunResult(tn, tim, wall, tn)


def _run_device_gravity(eng, prm, g, max_steps, t_end, dt_out, t_out,
                        start=(0, 0.0)) -> RunResult:
    damp = math.exp(-prm.alpha9wave * prm.courant) if (prm.mag_on and g.dim > 1) else 0.0
    tn, tim = int(start[0]), float(start[1])
    grvtime, cgrav_old = 0.0, 0.0
    if len(start) > 2:  # restart.f90:59
        grvtime, cgrav_old = tim, float(start[2])
    budget = max_steps if max_steps is not None else 1 << 62
    nsub_total = 0
    t0 = _time.perf_counter()
    while True:
        tim, tn, t_out, reason, grvtime, cgrav_old, ns = eng.run_loop_gravity(
            tim, tn, t_out, dt_out, t_end, prm.outstyle, prm.endstyle, prm.tnlim, prm.tn_out,
            budget - tn, damp, grvtime, cgrav_old)
        nsub_total += ns
        if reason != 2 or tn >= budget:
            break
    wall = _time.perf_counter() - t0
    res = RunResult(tn, tim, wall, tn, cgrav_old=cgrav_old)
    res.nsub_total = nsub_total
    return res


def fetch(eng, g: Grid, fields) -> dict:
    out = {}
    for f in fields:
        a = np.zeros(g.padded_shape)
        eng.download(f, a)
        out[f] = Grid.interior(a).copy()
    return out


SNAP_FIELDS = (F_D, F_V1, F_V2, F_V3, F_E)
SNAP_FIELDS_MAG = SNAP_FIELDS + (F_B1, F_B2, F_B3, F_PHI)


def snapshot(eng, g: Grid, prm: P.Params, res: RunResult, species=()) -> dict:
    """The state write_bin writes (output.f90:697-795) at the end of run
    result `res`: for snapshot.write_bin and load_restart_state."""
    names = ("d", "v1", "v2", "v3", "e", "b1", "b2", "b3", "phi")
    fields = SNAP_FIELDS_MAG if prm.mag_on else SNAP_FIELDS
    out = fetch(eng, g, fields)
    snap = {names[q]: out[f] for q, f in enumerate(fields)}
    snap["tn"], snap["time"] = res.tn, res.time
    if prm.gravswitch == 3:
        gv = fetch(eng, g, (F_GRVPHI, F_GRVPSI))
        snap["grvphi"], snap["grvpsi"] = gv[F_GRVPHI], gv[F_GRVPSI]
        snap["cgrav_old"] = res.cgrav_old
    if prm.compswitch == 2:
        sp = fetch(eng, g, tuple(F_SPC + q for q in range(prm.spn)))
        for q in range(prm.spn):
            snap[f"spc{q + 1}"] = sp[F_SPC + q]
        snap["species"] = list(species) or [f"spc{q + 1}" for q in range(prm.spn)]
    return snap
