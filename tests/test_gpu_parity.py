"""HIP path vs the reference: bit-exact FP64 parity on the GPU.

Each case runs the same initial state through the C-ABI (hg_*) and compares
the final state with the reference's own golden snapshot (or a reference run
at the benchmark size) byte for byte, and with the CPU oracle on the same
inputs.
"""
import ctypes

import numpy as np
import pytest

from hormone_amd.native import Hydro
from oracle.pyoracle import Oracle, lib as oracle_lib
from tests.helpers import (EXTRA_CASES, GOLDEN, compare_to_meta, compare_to_npz, digest,
                           meta, run_extra, run_problem)

pytestmark = pytest.mark.gpu

ONE_D = [f"{k}_{d}" for k in ("sodshock", "briowushock") for d in "xyz"]


def _kat(hiplib, which, rows, nout):
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    out = np.zeros((rows.shape[0], nout))
    dp = ctypes.POINTER(ctypes.c_double)
    hiplib.hg_kat.argtypes = [ctypes.c_int32, ctypes.c_int32, dp, dp]
    st = hiplib.hg_kat(which, rows.shape[0], rows.ctypes.data_as(dp), out.ctypes.data_as(dp))
    assert st == 0
    return out


def _random_riemann(n, rng):
    """Left/right states covering sub/supersonic flow, B_n = 0, psi != 0."""
    dl, dr = rng.lognormal(0, 1, n), rng.lognormal(0, 1, n)
    pl, pr = rng.lognormal(0, 1, n), rng.lognormal(0, 1, n)
    v = rng.uniform(-3, 3, (6, n))
    b = rng.uniform(-1, 1, (6, n))
    b[0:2, : n // 8] = 0.0  # B_n = 0
    phi = rng.uniform(-0.5, 0.5, (2, n))
    gam = 5.0 / 3.0
    cs_l, cs_r = np.sqrt(gam * pl / dl), np.sqrt(gam * pr / dr)
    ol = oracle_lib()
    cfl = np.array([ol.or_kat_get_cf(dl[i], cs_l[i], b[0, i], b[2, i], b[4, i]) for i in range(n)])
    cfr = np.array([ol.or_kat_get_cf(dr[i], cs_r[i], b[1, i], b[3, i], b[5, i]) for i in range(n)])
    el = pl / (gam - 1) + 0.5 * dl * (v[0] ** 2 + v[2] ** 2 + v[4] ** 2) + 0.5 * (b[0] ** 2 + b[2] ** 2 + b[4] ** 2)
    er = pr / (gam - 1) + 0.5 * dr * (v[1] ** 2 + v[3] ** 2 + v[5] ** 2) + 0.5 * (b[1] ** 2 + b[3] ** 2 + b[5] ** 2)
    ptl = pl + 0.5 * (b[0] ** 2 + b[2] ** 2 + b[4] ** 2)
    ptr = pr + 0.5 * (b[1] ** 2 + b[3] ** 2 + b[5] ** 2)
    ch = np.maximum(cfl, cfr) + 1.0
    rows = np.stack([cfl, cfr, v[0], v[1], v[2], v[3], v[4], v[5], dl, dr, el, er, ptl, ptr,
                     b[0], b[1], b[2], b[3], b[4], b[5], phi[0], phi[1], ch,
                     np.zeros(n), np.zeros(n)], axis=1)
    return rows


def test_kat_hlld_bitexact_vs_oracle(hiplib):
    rng = np.random.default_rng(20261202)
    rows = _random_riemann(4096, rng)
    gpu = _kat(hiplib, 0, rows, 9)
    ol = oracle_lib()
    dp = ctypes.POINTER(ctypes.c_double)
    ref = np.zeros((rows.shape[0], 10))
    for i in range(rows.shape[0]):
        r = np.ascontiguousarray(rows[i])
        ol.or_kat_hlldflux(r.ctypes.data_as(dp), ref[i].ctypes.data_as(dp))
    assert np.array_equal(gpu.view(np.uint64), ref[:, :9].view(np.uint64))


def test_kat_div_sqrt_correctly_rounded(hiplib):
    rng = np.random.default_rng(7)
    a = np.exp(rng.uniform(-300, 300, 20000))
    b = np.exp(rng.uniform(-300, 300, 20000)) * rng.choice([-1, 1], 20000)
    out = _kat(hiplib, 4, np.stack([a, b], 1), 2)
    assert np.array_equal(out[:, 0].view(np.uint64), (a / b).view(np.uint64))
    assert np.array_equal(out[:, 1].view(np.uint64), np.sqrt(a).view(np.uint64))


def _division_operands(rng, n):
    """Denominators and numerators over the whole double range: ordinary
    values, zeros of both signs, denormals, values next to the overflow and
    denormal thresholds, infinities and NaNs, and quotients that overflow or
    underflow -- the operands for which v_div_scale rescales."""
    special = np.array([0.0, -0.0, 5e-324, -5e-324, 1e-310, 2.0 ** -1022, 2.0 ** -1000,
                        2.0 ** -970, 2.0 ** -960, 1.0, -1.0, 3.0, 2.0 ** 767, 2.0 ** 1000,
                        2.0 ** 1022, 1.7e308, np.inf, -np.inf, np.nan])
    wide = np.exp(rng.uniform(-740, 709, n)) * rng.choice([-1.0, 1.0], n)
    b = np.where(rng.random(n) < 0.15, rng.choice(special, n), wide)
    a = np.exp(rng.uniform(-740, 709, (n, 4))) * rng.choice([-1.0, 1.0], (n, 4))
    pick = rng.random((n, 4))
    a = np.where(pick < 0.15, rng.choice(special, (n, 4)), a)
    with np.errstate(all="ignore"):
        near = b[:, None] * np.exp2(rng.choice([-1080.0, -1060.0, -1030.0, 760.0, 800.0, 1100.0],
                                               (n, 4)))
    a = np.where((pick >= 0.15) & (pick < 0.25), near, a)
    a = np.where((pick >= 0.25) & (pick < 0.35), rng.uniform(-4, 4, (n, 4)), a)
    return b, a


def test_kat_shared_denominator_division_is_plain_division(hiplib):
    """The sweep divides several numerators by one denominator with a shared
    reciprocal; every quotient must equal the correctly rounded a / b."""
    rng = np.random.default_rng(11)
    b, a = _division_operands(rng, 200_000)
    out = _kat(hiplib, 5, np.concatenate([b[:, None], a], 1), 4)
    with np.errstate(all="ignore"):
        ref = a / b[:, None]
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(out), nan)
    assert np.array_equal(out[~nan].view(np.uint64), ref[~nan].view(np.uint64))


@pytest.mark.parametrize("flat", [False, True], ids=["muscl", "flat"])
@pytest.mark.parametrize("name", ONE_D)
def test_gpu_1d_goldens(hiplib, name, flat):
    print(f"case {name} flat={flat}", flush=True)
    res, state, _ = run_problem(Hydro, name, flat=flat)
    compare_to_npz(res, state, name + ("_flat" if flat else ""))
    # and the reference's own test criterion (tests.f90:48-178) says Identical
    from hormone_amd import problems as P
    from hormone_amd import testnorms as T
    prm, g, _ = P.make_problem(name)
    z = np.load(f"{GOLDEN}/{name}{'_flat' if flat else ''}.npz")
    golden = {k: z[k].reshape(g.shape) for k in state}
    mine = {k: v.reshape(g.shape) for k, v in state.items()}
    passed, errs, identical = T.reference_criterion(
        mine, golden, prm.gamma, T.cell_volumes(g), T.tolerance_for(name), prm.mag_on, g.dim)
    assert passed and identical, errs


@pytest.mark.parametrize("flat", [False, True], ids=["muscl", "flat"])
@pytest.mark.parametrize("plane", ["xy", "xz", "yz"])
def test_gpu_orszagtang_goldens(hiplib, plane, flat):
    print(f"case orszagtang_{plane} flat={flat}", flush=True)
    res, state, _ = run_problem(Hydro, f"orszagtang_{plane}", flat=flat)
    compare_to_meta(res, state, meta()[f"orszagtang_{plane}" + ("_flat" if flat else "")])


@pytest.mark.parametrize("flat", [False, True], ids=["muscl", "flat"])
def test_gpu_kh_goldens(hiplib, flat):
    """Kelvin-Helmholtz with two advected species (compswitch 2): the
    reference's KHinstability goldens bit for bit, species included."""
    print(f"case KHinstability flat={flat}", flush=True)
    res, state, _ = run_problem(Hydro, "KHinstability", flat=flat)
    compare_to_meta(res, state, meta()["KHinstability" + ("_flat" if flat else "")])


def test_gpu_sod1024_courant04(hiplib):
    res, state, _ = run_problem(Hydro, "sodshock_x", n=1024, courant=0.4)
    compare_to_npz(res, state, "sod1024_c04")


def test_gpu_ot512_prefix(hiplib):
    res, state, _ = run_problem(Hydro, "orszagtang_xy", n=512, outstyle=2, endstyle=2,
                                tnlim=20)
    compare_to_meta(res, state, meta()["ot512_20steps"])


def test_gpu_matches_oracle_midrun(hiplib):
    """Same inputs, 37 steps of OT 96^2: every field equal to the oracle."""
    r1, s1, _ = run_problem(Hydro, "orszagtang_xy", n=96, max_steps=37)
    r2, s2, _ = run_problem(Oracle, "orszagtang_xy", n=96, max_steps=37)
    assert r1.time == r2.time
    for k in s2:
        assert np.array_equal(s1[k].view(np.uint64), s2[k].view(np.uint64)), k


@pytest.mark.parametrize("name", ["sodshock_x", "briowushock_y", "orszagtang_xy", "orszagtang_yz"])
def test_gpu_device_loop_goldens(hiplib, name):
    """The device-resident main loop (hg_run) reproduces the goldens too."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    print(f"case device-loop {name}", flush=True)
    prm, g, st = P.make_problem(name)
    eng = Hydro(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    res = D.run(eng, prm, g, device_loop=True)
    fields = D.SNAP_FIELDS_MAG if prm.mag_on else D.SNAP_FIELDS
    out = D.fetch(eng, g, fields)
    names = ("d", "v1", "v2", "v3", "e", "b1", "b2", "b3", "phi")
    state = {names[i]: out[f] for i, f in enumerate(fields)}
    if name.startswith("orszagtang"):
        compare_to_meta(res, state, meta()[name])
    else:
        compare_to_npz(res, state, name)


def _ot512_engine():
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    prm, g, st = P.make_problem("orszagtang_xy", n=512)
    eng = Hydro(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    return prm, g, eng


def test_gpu_ot512_2200_steps_and_reference_stop(hiplib):
    """The benchmark workload: state after 2200 steps equals the reference
    program's, and the run then stops where the reference stops (negative
    internal energy at the same cell, same step)."""
    import math
    from hormone_amd import driver as D
    from hormone_amd.native import HG_ERR_NEG_EINT, HgError
    m = meta()
    if "ot512_2200steps" not in m or "ot512_t05_failure" not in m:
        pytest.skip("reference-run fixtures not generated")
    prm, g, eng = _ot512_engine()
    t_end, dt_out, t_out = prm.times()
    damp = math.exp(-prm.alpha9wave * prm.courant)
    tim, tn, t_out, reason = eng.run_loop(0.0, 0, t_out, dt_out, t_end, 1, 1, 0, 1, 2200, damp)
    assert tn == 2200
    out = D.fetch(eng, g, D.SNAP_FIELDS_MAG)
    names = ("d", "v1", "v2", "v3", "e", "b1", "b2", "b3", "phi")
    state = {names[q]: out[f] for q, f in enumerate(D.SNAP_FIELDS_MAG)}
    e = m["ot512_2200steps"]
    assert tim == e["time"]
    for nm in names:
        assert digest(state[nm]) == e["fields"][nm]["sha256"], nm
    fail = m["ot512_t05_failure"]
    with pytest.raises(HgError) as ei:
        eng.run_loop(tim, tn, t_out, dt_out, t_end, 1, 1, 0, 1, 1000, damp)
    assert ei.value.code == HG_ERR_NEG_EINT
    assert ei.value.tn == fail["completed_steps"]
    assert list(ei.value.cell) == fail["cell"]


def test_gpu_gravity_laplacian_quadratic(hiplib):
    from tests.test_gravity import lap_from_one_substep, quadratic_setup
    prm, g, eng = quadratic_setup(Hydro)
    lap = lap_from_one_substep(eng, g, prm)
    assert np.max(np.abs(lap - 12.0)) < 1e-9


@pytest.mark.parametrize("mag_on,n,steps",
                         [(False, 32, 12), (True, 32, 12), (True, 64, 4), (True, 128, 2)],
                         ids=["hydro32", "mhd32", "mhd64", "mhd128"])
def test_gpu_polytrope_gravity_matches_oracle(hiplib, mag_on, n, steps):
    """3D hydro (or MHD with B=0) + hyperbolic gravity: every field and the
    potential equal the oracle bit for bit after 12 steps at 32^3 (gravity
    blocks of 64 x 4 cells), 4 steps at 64^3 (one 64-cell row per block)
    and 2 steps at 128^3 (one row per block, two cells per thread)."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    print(f"case polytrope{n} mag_on={mag_on}", flush=True)
    res = []
    for cls in (Hydro, Oracle):
        prm, g, st = P.make_problem("polytrope3d", n=n, steps=steps, mag_on=mag_on)
        eng = cls(**D.desc_fields(prm, g))
        D.load_initial_state(eng, prm, g, st)
        r = D.run(eng, prm, g)
        flds = (D.SNAP_FIELDS_MAG if mag_on else D.SNAP_FIELDS) + (D.F_GRVPHI, D.F_GRVPSI)
        res.append((r, D.fetch(eng, g, flds)))
    (r1, s1), (r2, s2) = res
    assert r1.tn == r2.tn == steps and r1.time == r2.time and r1.nsub == r2.nsub
    for f in s2:
        assert np.array_equal(s1[f].view(np.uint64), s2[f].view(np.uint64)), f


@pytest.mark.parametrize("engine", ["hip", "oracle"])
def test_negative_internal_energy_is_reported_at_its_cell(hiplib, engine):
    """A state whose total energy is below its kinetic energy in one cell
    stops like internalenergy (eos.f90:833-839), naming that cell."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    from hormone_amd.native import HG_ERR_NEG_EINT, HgError
    prm, g, st = P.make_problem("sodshock_x", n=100)
    st.v[0][0, 0, 36] = 3.0  # Fortran cell i = 37
    eint, e = P.initial_energy(st, prm.gamma)
    e[0, 0, 36] = 0.25 * st.d[0, 0, 36] * 9.0  # below the kinetic energy 0.5*d*v^2
    u = P.conserve(st, e, prm.mag_on, g.dim)
    eng = (Hydro if engine == "hip" else Oracle)(**D.desc_fields(prm, g))
    for f, a in {D.F_D: st.d, D.F_V1: st.v[0], D.F_V2: st.v[1], D.F_V3: st.v[2], D.F_E: e}.items():
        eng.upload(f, g.pad(a))
    for n, a in enumerate(u):
        eng.upload(D.F_U + n, g.pad(a))
    with pytest.raises(HgError) as ei:
        eng.timestep(0.0, 0.2, 0.2, 1, 1)
    assert ei.value.code == HG_ERR_NEG_EINT
    assert tuple(ei.value.cell) == (37, 1, 1)


def test_large_contexts_created_back_to_back_start_clean(hiplib):
    """Regression: context set-up zeroes its buffers on its own stream before
    clearing the error key.  Zeroing on the null stream was not ordered with
    the context's non-blocking stream, so a second large context (kernels
    already loaded, clear issued at once) could read a spurious error."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    engs = []
    for q in range(2):
        print(f"case back-to-back context {q}", flush=True)
        prm, g, st = P.make_problem("kh3d", n=160, steps=2)
        eng = Hydro(**D.desc_fields(prm, g))
        D.load_initial_state(eng, prm, g, st)
        r = D.run(eng, prm, g, device_loop=True)
        assert r.tn == 2
        engs.append(eng)


@pytest.mark.parametrize("case", list(EXTRA_CASES))
def test_gpu_extra_reference_runs(hiplib, case):
    """Reflective and outgoing boundaries, minmod, RK2 and RK1: bit-identical
    to runs of the reference program (tests/golden/make_fixtures.py)."""
    print(f"case extra {case}", flush=True)
    res, state, _ = run_extra(Hydro, case)
    compare_to_meta(res, state, meta()[case])


def test_gpu_reflective_briowu_stops_like_the_reference(hiplib):
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    from hormone_amd.native import HG_ERR_NEG_EINT, HgError
    fail = meta()["briowu_x_reflective_failure"]
    name, kw = EXTRA_CASES["briowu_x_reflective_5steps"]
    kw = dict(kw, outstyle=1, endstyle=1, t_end=300.0, dt_out=300.0)
    prm, g, st = P.make_problem(name, **kw)
    eng = Hydro(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    res = D.run(eng, prm, g, max_steps=fail["completed_steps"])
    assert res.tn == fail["completed_steps"]
    with pytest.raises(HgError) as ei:
        D.run(eng, prm, g, max_steps=1)
    assert ei.value.code == HG_ERR_NEG_EINT
    assert list(ei.value.cell) == fail["cell"]


def test_gpu_linear_boundary_stops_like_the_reference(hiplib):
    """Orszag-Tang yz with linear codes at the x3 boundaries stops during
    step 50 at the reference's cell (negative internal energy)."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    from hormone_amd.native import HG_ERR_NEG_EINT, HgError
    print("case ot_yz_linear_failure", flush=True)
    fail = meta()["ot_yz_linear_failure"]
    name, kw = EXTRA_CASES["ot_yz_linear"]
    prm, g, st = P.make_problem(name, **dict(kw, tnlim=100, tn_out=100))
    eng = Hydro(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    res = D.run(eng, prm, g, max_steps=fail["completed_steps"])
    assert res.tn == fail["completed_steps"]
    with pytest.raises(HgError) as ei:
        D.run(eng, prm, g, max_steps=1)
    assert ei.value.code == HG_ERR_NEG_EINT
    assert list(ei.value.cell) == fail["cell"]


def test_time_sweep_leaves_the_state_unchanged(hiplib):
    """hg_time_sweep (the benchmark's kernel timing) replays sweeps between
    steps: the run continues bit for bit as if it had not been called, also
    after a replay on a run that has reached its end."""
    import bench as B
    from hormone_amd import driver as D
    from hormone_amd import problems as P

    def start():
        prm, g, st = P.make_problem("orszagtang_xy", n=64, outstyle=2, endstyle=2,
                                    tnlim=30, tn_out=30)
        eng = Hydro(**D.desc_fields(prm, g))
        D.load_initial_state(eng, prm, g, st)
        return g, eng, B.Runner(eng, prm, g)

    g, e1, r1 = start()
    r1.advance(10)
    assert e1.time_sweep(7) > 0.0
    r1.advance(20)
    e1.time_sweep(3)  # at the end of the run (stop flag set)
    g, e2, r2 = start()
    r2.advance(30)
    assert (r1.tn, r1.tim) == (r2.tn, r2.tim)
    a = D.fetch(e1, g, D.SNAP_FIELDS_MAG)
    b = D.fetch(e2, g, D.SNAP_FIELDS_MAG)
    for f in D.SNAP_FIELDS_MAG:
        assert np.array_equal(a[f].view(np.uint64), b[f].view(np.uint64)), f


@pytest.mark.parametrize("bw", [4, 8, 16])
@pytest.mark.parametrize("name,kw,steps", [
    ("orszagtang_xy", dict(n=90), 15),
    ("orszagtang_yz", dict(n=70), 10),
    ("kh3d", dict(n=28, steps=6), 6),
])
def test_gpu_tiled_sweep_any_width_matches_oracle(hiplib, name, kw, steps, bw):
    """The LDS tile of the y/z sweeps with x extents that are not a multiple
    of the block's line count (blocks spanning the end of an x row or of a
    segment, padding wavefronts in the last block): bitwise the oracle."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    print(f"case tiled {name} {kw} bw={bw}", flush=True)
    prm, g, st = P.make_problem(name, **kw)
    e1 = Hydro(**D.desc_fields(prm, g))
    e1.set_launch(bw)
    D.load_initial_state(e1, prm, g, st)
    r1 = D.run(e1, prm, g, max_steps=steps)
    e2 = Oracle(**D.desc_fields(prm, g))
    D.load_initial_state(e2, prm, g, st)
    r2 = D.run(e2, prm, g, max_steps=steps)
    assert (r1.tn, r1.time) == (r2.tn, r2.time)
    fields = D.SNAP_FIELDS_MAG if prm.mag_on else D.SNAP_FIELDS
    a, b = D.fetch(e1, g, fields), D.fetch(e2, g, fields)
    for f in fields:
        assert np.array_equal(a[f].view(np.uint64), b[f].view(np.uint64)), f


@pytest.mark.parametrize("bw", [1, 4, 16])
@pytest.mark.parametrize("name,kw,steps,mode", [
    ("orszagtang_xy", dict(n=90), 15, 1),
    ("orszagtang_xy", dict(n=90), 15, 2),
    ("orszagtang_yz", dict(n=70), 10, 1),
    ("briowushock_y", dict(), 40, 1),
    ("sodshock_x", dict(), 40, 2),
    ("kh3d", dict(n=28, steps=6), 6, 1),
    ("kh3d", dict(n=28, steps=6), 6, 2),
])
def test_gpu_fused_last_sweep_matches_oracle(hiplib, name, kw, steps, mode, bw):
    """A direction's sweep fused with the RK update (in-place update from a
    halo copy; mode 1 the last direction, 2 x), lane lines (bw=1) and LDS
    tiles with padding columns: bitwise the oracle."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    print(f"case fused {name} {kw} mode={mode} bw={bw}", flush=True)
    prm, g, st = P.make_problem(name, **kw)
    e1 = Hydro(**D.desc_fields(prm, g))
    e1.set_launch(bw)
    e1.set_fused(mode)
    D.load_initial_state(e1, prm, g, st)
    r1 = D.run(e1, prm, g, max_steps=steps)
    e2 = Oracle(**D.desc_fields(prm, g))
    D.load_initial_state(e2, prm, g, st)
    r2 = D.run(e2, prm, g, max_steps=steps)
    assert (r1.tn, r1.time) == (r2.tn, r2.time)
    fields = D.SNAP_FIELDS_MAG if prm.mag_on else D.SNAP_FIELDS
    a, b = D.fetch(e1, g, fields), D.fetch(e2, g, fields)
    for f in fields:
        assert np.array_equal(a[f].view(np.uint64), b[f].view(np.uint64)), f


@pytest.mark.parametrize("bw", [1, 8])
@pytest.mark.parametrize("name,mode", [("briowushock_y", 1), ("orszagtang_xy", 1),
                                       ("orszagtang_yz", 1), ("orszagtang_xy", 2),
                                       ("sodshock_x", 2)])
def test_gpu_fused_device_loop_goldens(hiplib, name, mode, bw):
    """hg_run with the fused last sweep: the CFL partials come one per block
    of the fused launch, and the goldens still hold bit for bit."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    print(f"case fused device-loop {name} mode={mode} bw={bw}", flush=True)
    prm, g, st = P.make_problem(name)
    eng = Hydro(**D.desc_fields(prm, g))
    eng.set_launch(bw)
    eng.set_fused(mode)
    eng.profile(bw == 8)  # per-launch events around the fused launches too
    D.load_initial_state(eng, prm, g, st)
    res = D.run(eng, prm, g, device_loop=True)
    if bw == 8:
        from hormone_amd.native import PROF_RK
        ms, n, _ = eng.profile_read(PROF_RK)
        assert n == prm.rktype * res.tn and ms > 0.0
    fields = D.SNAP_FIELDS_MAG if prm.mag_on else D.SNAP_FIELDS
    out = D.fetch(eng, g, fields)
    names = ("d", "v1", "v2", "v3", "e", "b1", "b2", "b3", "phi")
    state = {names[i]: out[f] for i, f in enumerate(fields)}
    if name.startswith("orszagtang"):
        compare_to_meta(res, state, meta()[name])
    else:
        compare_to_npz(res, state, name)


@pytest.mark.parametrize("mode", [1, 2])
def test_gpu_fused_polytrope_gravity_matches_oracle(hiplib, mode):
    """Fused z (or x) sweep with the gravity source in the RK update: every
    field and the potential equal the oracle after 12 steps at 32^3."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    print(f"case fused polytrope32 mode={mode}", flush=True)
    res = []
    for cls in (Hydro, Oracle):
        prm, g, st = P.make_problem("polytrope3d", n=32, steps=12)
        eng = cls(**D.desc_fields(prm, g))
        if cls is Hydro:
            eng.set_fused(mode)
        D.load_initial_state(eng, prm, g, st)
        r = D.run(eng, prm, g)
        res.append((r, D.fetch(eng, g, D.SNAP_FIELDS + (D.F_GRVPHI, D.F_GRVPSI))))
    (r1, s1), (r2, s2) = res
    assert r1.tn == r2.tn == 12 and r1.time == r2.time
    for f in s2:
        assert np.array_equal(s1[f].view(np.uint64), s2[f].view(np.uint64)), f


def test_gpu_fused_rejects_what_it_does_not_cover(hiplib):
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    from hormone_amd.native import HG_ERR_UNSUPPORTED, HgError
    for name, mode in (("sodshock_x", 1), ("briowushock_y", 2), ("KHinstability", 1),
                       ("KHinstability", 2)):  # no y/z; no x; species
        prm, g, st = P.make_problem(name)
        eng = Hydro(**D.desc_fields(prm, g))
        with pytest.raises(HgError) as ei:
            eng.set_fused(mode)
        assert ei.value.code == HG_ERR_UNSUPPORTED
        eng.set_fused(False)


@pytest.mark.parametrize("axis", ["x", "z"])
@pytest.mark.parametrize("mag", [False, True], ids=["hydro", "mhd"])
def test_gpu_dirichlet_and_flux_boundary_closed_forms(hiplib, axis, mag):
    """Dirichlet ghosts holding a uniform flow keep it uniform bit for bit;
    flux boundaries change every conserved total by dt*(F_in - F_out)*area."""
    from hormone_amd import driver as D
    from tests import boundary_cases as B
    print(f"case boundary closed forms {axis} mag={mag}", flush=True)
    prm, g, eng = B.uniform(Hydro, axis, 9, mag=mag)
    B.set_uniform_ghosts(eng, g, mag)
    f0 = D.fetch(eng, g, B.fields(prm))
    D.run(eng, prm, g, max_steps=10)
    f1 = D.fetch(eng, g, B.fields(prm))
    for f in f0:
        assert np.array_equal(f0[f], f1[f]), f
    got, want = B.flux_budget(Hydro, axis, mag)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-15)


def _ext_case(kind):
    """A smooth 2D MHD state on the yz plane (64^2 cells of Orszag-Tang's
    grid, periodic y) with the extended codes on the z faces; the caller
    sets Dirichlet ghosts and boundary fluxes near that state."""
    from hormone_amd import problems as P
    bc = {"dirichlet": (0, 0, 0, 0, 9, 9, 0, 0, 0, 0, 9, 9),
          "flux": (0, 0, 0, 0, 10, 10, 0, 0, 0, 0, 10, 10),
          "mixed": (0, 0, 0, 0, 10, 4, 0, 0, 0, 0, 9, 5)}[kind]
    prm, g, st = P.make_problem("orszagtang_yz", n=64, bc=bc, outstyle=2, endstyle=2,
                                tnlim=1000, tn_out=1000)
    y = g.centres(1)[None, :, None]
    z = g.centres(2)[:, None, None]
    s = np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z)
    c = np.cos(2 * np.pi * y) * np.cos(2 * np.pi * z)
    st.d[...] = 1.0 + 0.1 * s
    st.p[...] = 0.6 + 0.05 * c
    for a in range(3):
        st.v[a][...] = 0.05 * (s if a != 1 else c)
        st.b[a][...] = 0.1 * (c if a != 2 else s) + (0.3 if a == 2 else 0.0)
    return prm, g, st


def _ext_boundary_values(eng, g, phase):
    from hormone_amd import driver as D
    rng = np.random.default_rng(7 + phase)
    sh = g.padded_shape
    eng.set_ghosts(D.F_D, 1.0 + 0.1 * rng.random(sh))
    eng.set_ghosts(D.F_P, 0.6 + 0.05 * rng.random(sh))
    for a in range(3):
        eng.set_ghosts(D.F_V1 + a, 0.05 * (rng.random(sh) - 0.5))
        eng.set_ghosts(D.F_B1 + a, 0.1 * (rng.random(sh) - 0.5) + (0.3 if a == 2 else 0.0))
    eng.set_ghosts(D.F_PHI, 0.01 * rng.random(sh))
    # z faces: normal momentum flux (imo3) near the total pressure
    flux = np.array([0.001, 0.0, 0.002, 0.6, 0.01, 0.0, 0.001, -0.001, 0.0]) * (1 + 0.1 * phase)
    for face in (4, 5):
        try:
            eng.set_boundary_flux(face, flux * (1.0 if face == 4 else 0.5))
        except Exception:  # not a flux face
            pass


@pytest.mark.parametrize("bw", [1, 8])
@pytest.mark.parametrize("kind", ["dirichlet", "flux", "mixed"])
def test_gpu_extended_boundaries_match_oracle(hiplib, kind, bw):
    """Dirichlet, flux and linear codes on the z faces, lane lines (bw=1) and
    LDS tiles: bitwise the oracle, with the boundary data changed between
    steps (time-dependent Dirichlet data)."""
    from hormone_amd import driver as D
    print(f"case extended boundaries {kind} bw={bw}", flush=True)
    out = []
    for cls in (Hydro, Oracle):
        prm, g, st = _ext_case(kind)
        eng = cls(**D.desc_fields(prm, g))
        if cls is Hydro:
            eng.set_launch(bw)
        D.load_initial_state(eng, prm, g, st)
        _ext_boundary_values(eng, g, 0)
        r1 = D.run(eng, prm, g, max_steps=8)
        _ext_boundary_values(eng, g, 1)
        r2 = D.run(eng, prm, g, max_steps=16, start=(r1.tn, r1.time))
        out.append((r2.tn, r2.time, D.fetch(eng, g, D.SNAP_FIELDS_MAG)))
    (n1, t1, a), (n2, t2, b) = out
    assert (n1, t1) == (n2, t2) == (16, t2)
    for f in D.SNAP_FIELDS_MAG:
        assert np.array_equal(a[f].view(np.uint64), b[f].view(np.uint64)), f


@pytest.mark.parametrize("bw", [1, 16])
def test_gpu_linear_boundaries_any_launch_shape(hiplib, bw):
    """The linear codes (and the unwritten v2/b2 ghost below the x3 inner
    boundary) in lane lines and in 16-line tiles: the reference's run."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    print(f"case ot_yz_linear bw={bw}", flush=True)
    name, kw = EXTRA_CASES["ot_yz_linear"]
    prm, g, st = P.make_problem(name, **kw)
    eng = Hydro(**D.desc_fields(prm, g))
    eng.set_launch(bw)
    D.load_initial_state(eng, prm, g, st)
    res = D.run(eng, prm, g)
    out = D.fetch(eng, g, D.SNAP_FIELDS_MAG)
    names = ("d", "v1", "v2", "v3", "e", "b1", "b2", "b3", "phi")
    compare_to_meta(res, {names[i]: out[f] for i, f in enumerate(D.SNAP_FIELDS_MAG)},
                    meta()["ot_yz_linear"])


def test_gpu_species_with_linear_boundary_matches_oracle(hiplib):
    """Advected species next to a linear (code 4) x1 outer face, whose
    species ghosts copy the edge cell (boundarycondition.f90:317-320):
    bitwise the oracle."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    print("case KH species + linear x1 outer", flush=True)
    bc = (2, 4, 0, 0, 0, 0, 2, 4, 0, 0, 0, 0)
    out = []
    for cls in (Hydro, Oracle):
        prm, g, st = P.make_problem("KHinstability", bc=bc)
        eng = cls(**D.desc_fields(prm, g))
        D.load_initial_state(eng, prm, g, st)
        r = D.run(eng, prm, g, max_steps=30)
        flds = D.SNAP_FIELDS + tuple(D.F_SPC + q for q in range(prm.spn))
        out.append((r.tn, r.time, D.fetch(eng, g, flds)))
    (n1, t1, a), (n2, t2, b) = out
    assert (n1, t1) == (n2, t2)
    for f in b:
        assert np.array_equal(a[f].view(np.uint64), b[f].view(np.uint64)), f


@pytest.mark.parametrize("mode", [1, 2])
def test_gpu_fused_device_loop_across_output_stops(hiplib, mode):
    """hg_run stops at every output time (outstyle 1) with steps already
    enqueued behind the stop; the fused form swaps its two state buffers at
    each of those stages too, so they must hand the state over unchanged.
    Many stops, then the end: bit for bit the separate RK update."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    print(f"case fused device loop output stops mode={mode}", flush=True)
    out = []
    for fused in (0, mode):
        prm, g, st = P.make_problem("orszagtang_xy", n=64, t_end=100.0, dt_out=7.0)
        eng = Hydro(**D.desc_fields(prm, g))
        eng.set_fused(fused)
        D.load_initial_state(eng, prm, g, st)
        res = D.run(eng, prm, g, device_loop=True)
        out.append((res.tn, res.time, D.fetch(eng, g, D.SNAP_FIELDS_MAG)))
    (n1, t1, a), (n2, t2, b) = out
    assert (n1, t1) == (n2, t2) and n1 > 20
    for f in a:
        assert np.array_equal(a[f].view(np.uint64), b[f].view(np.uint64)), f
