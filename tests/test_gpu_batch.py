"""Batched boxes and continuous-batching serving on the GPU.

Every box in a batch keeps its own dt and clock, so its final state must be
bit-identical to running it alone (hg_run), whatever else shares the
launches and whatever launch shapes the autotuner picks.
"""
import pytest

from hormone_amd import driver as D
from hormone_amd import serving as S
from hormone_amd.batch import Batch, autotune
from hormone_amd.native import Hydro

pytestmark = pytest.mark.gpu


def small_requests(count, seed=11):
    return S.make_requests(count, rate=50.0, seed=seed, min_cells=24 ** 2, max_cells=36 ** 3)


def test_batch_members_equal_alone(hiplib):
    reqs = small_requests(5)
    alone = [S.run_alone(r) for r in reqs]
    b = Batch()
    engs = []
    for r in reqs:
        prm, g, st = S.setup(r)
        eng = Hydro(**D.desc_fields(prm, g))
        D.load_initial_state(eng, prm, g, st)
        t_end, dt_out, t_out = prm.times()
        b.add(eng, t_end=t_end, t_out=t_out, dt_out=dt_out)
        engs.append((eng, g))
    while b.step(8):
        pass
    for (eng, g), r, (tn, tm, dg) in zip(engs, reqs, alone):
        t, n, stop = b.state(eng)
        assert stop == 1 and n == tn and t == tm, (r.rid, n, tn)
        b.remove(eng)
        assert S.state_digest(eng, g) == dg, f"request {r.rid} differs in the batch"


def test_batch_launch_shapes_do_not_change_results(hiplib):
    reqs = small_requests(3, seed=5)
    digests = []
    # (4, 64, 0): a 64-block persistent grid, so every tiled block runs many
    # chunks in a row (the barrier between chunks guards the tile's table)
    for shape in [(4, 0, 0), (1, 512, 1024), (2, 2048, 0), (16, 1024, 0), (8, 0, 2048),
                  (4, 64, 0)]:
        b = Batch()
        b.configure(*shape, use_graph=shape[0] != 1)
        engs = []
        for r in reqs:
            prm, g, st = S.setup(r)
            eng = Hydro(**D.desc_fields(prm, g))
            D.load_initial_state(eng, prm, g, st)
            t_end, dt_out, t_out = prm.times()
            b.add(eng, t_end=t_end, t_out=t_out, dt_out=dt_out)
            engs.append((eng, g))
        while b.step(16):
            pass
        digests.append([S.state_digest(e, g) for e, g in engs])
        b.close()
    assert all(d == digests[0] for d in digests[1:])


def test_autotune_picks_a_shape(hiplib):
    reqs = small_requests(4, seed=3)
    b = Batch()
    for r in reqs:
        prm, g, st = S.setup(r)
        eng = Hydro(**D.desc_fields(prm, g))
        D.load_initial_state(eng, prm, g, st)
        t_end, dt_out, t_out = prm.times()
        b.add(eng, t_end=t_end, t_out=t_out, dt_out=dt_out)
    best, timings = autotune(b, steps_per_trial=1, use_cache=False)
    assert best in timings and len(timings) >= 4


@pytest.mark.parametrize("lanes", [(), (2000,)], ids=["one-batch", "two-size-classes"])
def test_serving_results_equal_alone_and_report(hiplib, lanes):
    """Every request ends as it does alone, whether it shares one batch with
    all others or the size classes run concurrently on their own streams."""
    reqs = small_requests(10, seed=21)
    rep = S.serve(reqs, hbm_budget=2 ** 30, steps_per_round=4, tune=True, keep_digest=True,
                  lanes=lanes)
    summ = rep.summary()
    assert summ["completed"] == 10 and summ["failed"] == 0
    assert summ["p50_s"] is not None and summ["p99_s"] >= summ["p50_s"]
    for r in reqs[:4]:
        tn, tm, dg = S.run_alone(r)
        assert r.steps == tn and r.digest == dg, f"request {r.rid} differs in the mix"


def test_batch_remove_and_state_before_any_step(hiplib):
    """A member removed, or asked for its state, before the batch has
    stepped: no stale host record is read (HG_ERR_STATE for the state) and
    the members that stay run on to their alone result."""
    from hormone_amd.native import HG_ERR_STATE, HgError
    reqs = small_requests(3, seed=31)
    b = Batch()
    engs = []
    for r in reqs:
        prm, g, st = S.setup(r)
        eng = Hydro(**D.desc_fields(prm, g))
        D.load_initial_state(eng, prm, g, st)
        t_end, dt_out, t_out = prm.times()
        b.add(eng, t_end=t_end, t_out=t_out, dt_out=dt_out)
        engs.append((eng, g, r))
    with pytest.raises(HgError) as ei:
        b.state(engs[0][0])
    assert ei.value.code == HG_ERR_STATE
    b.remove(engs[0][0])
    b.step(2)
    # added after the last step: no record yet
    prm, g, st = S.setup(reqs[0])
    late = Hydro(**D.desc_fields(prm, g))
    D.load_initial_state(late, prm, g, st)
    t_end, dt_out, t_out = prm.times()
    b.add(late, t_end=t_end, t_out=t_out, dt_out=dt_out)
    with pytest.raises(HgError) as ei:
        b.state(late)
    assert ei.value.code == HG_ERR_STATE
    b.remove(late)
    while b.step(8):
        pass
    for eng, g, r in engs[1:]:
        t, n, stop = b.state(eng)
        tn, tm, dg = S.run_alone(r)
        assert stop == 1 and n == tn and t == tm
        b.remove(eng)
        assert S.state_digest(eng, g) == dg


# Golden problems of the reference that a batch can hold (no species, no
# output stop before the end): one batch per kernel variant, members of
# mixed dimension and size.
GOLDEN_BATCHES = {
    "hydro-muscl": [("sodshock_x", {}, "sodshock_x"), ("sodshock_y", {}, "sodshock_y"),
                    ("sodshock_z", {}, "sodshock_z"),
                    ("sodshock_x", dict(n=1024, courant=0.4), "sod1024_c04")],
    "hydro-flat": [("sodshock_x", {}, "sodshock_x_flat"), ("sodshock_y", {}, "sodshock_y_flat"),
                   ("sodshock_z", {}, "sodshock_z_flat")],
    "mhd-muscl": [("briowushock_x", {}, "briowushock_x"), ("briowushock_y", {}, "briowushock_y"),
                  ("briowushock_z", {}, "briowushock_z"), ("orszagtang_xy", {}, "orszagtang_xy"),
                  ("orszagtang_xz", {}, "orszagtang_xz"), ("orszagtang_yz", {}, "orszagtang_yz")],
    "mhd-flat": [("briowushock_y", {}, "briowushock_y_flat"),
                 ("orszagtang_xy", {}, "orszagtang_xy_flat"),
                 ("orszagtang_yz", {}, "orszagtang_yz_flat")],
    # every member sweeps x: the batch runs in the x-fused form
    "hydro-muscl-x": [("sodshock_x", {}, "sodshock_x"),
                      ("sodshock_x", dict(n=1024, courant=0.4), "sod1024_c04")],
    "mhd-muscl-x": [("briowushock_x", {}, "briowushock_x"), ("orszagtang_xy", {}, "orszagtang_xy"),
                    ("orszagtang_xz", {}, "orszagtang_xz")],
    "mhd-flat-x": [("orszagtang_xy", {}, "orszagtang_xy_flat")],
}


@pytest.mark.parametrize("variant", list(GOLDEN_BATCHES))
def test_batch_of_golden_problems_matches_the_reference(hiplib, variant):
    """The reference's own golden snapshots, run as members of one batch:
    every member ends at the golden's step count and time with every field
    equal to the reference's bit for bit."""
    from types import SimpleNamespace

    from hormone_amd import problems as P
    from tests.helpers import NAMES, compare_to_meta, compare_to_npz, meta

    print(f"case batch of goldens {variant}", flush=True)
    flat = "flat" in variant
    b = Batch()
    members = []
    for name, kw, key in GOLDEN_BATCHES[variant]:
        prm, g, st = P.make_problem(name, flux_limiter="flat" if flat else "modified_mc", **kw)
        eng = Hydro(**D.desc_fields(prm, g))
        D.load_initial_state(eng, prm, g, st)
        t_end, dt_out, t_out = prm.times()
        assert t_out == t_end  # no output stop changes dt before the end
        b.add(eng, t_end=t_end, t_out=t_out, dt_out=dt_out, outstyle=prm.outstyle,
              endstyle=prm.endstyle, tnlim=prm.tnlim, tn_out=prm.tn_out)
        members.append((eng, g, prm, key))
    while b.step(64):
        pass
    # the x-fused form exactly when every member sweeps x
    assert b.fused() == all(g.ax[0].n > 1 for _, g, _, _ in members)
    m = meta()
    for eng, g, prm, key in members:
        t, n, stop = b.state(eng)
        assert stop == 1, (key, stop)
        b.remove(eng)
        fields = D.SNAP_FIELDS_MAG if prm.mag_on else D.SNAP_FIELDS
        out = D.fetch(eng, g, fields)
        state = {NAMES[i]: out[f] for i, f in enumerate(fields)}
        res = SimpleNamespace(tn=n, time=t)
        if key in m and "fields" in m[key] and "sha256" in next(iter(m[key]["fields"].values())):
            compare_to_meta(res, state, m[key])
        else:
            compare_to_npz(res, state, key)
    b.close()


def test_serving_ensemble_at_full_size_range(hiplib):
    """The serving ensemble over its full size range: a seeded draw of the
    default generator (64^2 .. 192^3 cells) whose first request is the
    smallest box (64^2) and whose last the largest (192^3).  Every request
    served in the mix ends as it does alone (bit for bit, or with the same
    error at the same step), and the report carries p50/p99."""
    from hormone_amd.native import HgError

    reqs = S.make_requests(6, rate=50.0, seed=2024)
    reqs[0].dim, reqs[0].n = 2, 64
    reqs[-1].dim, reqs[-1].n = 3, 192
    cells = [r.cells for r in reqs]
    assert min(cells) == 64 ** 2 and max(cells) == 192 ** 3
    print("case serving full range, cells", cells, flush=True)
    rep = S.serve(reqs, hbm_budget=64 * 2 ** 30, steps_per_round=4, tune=True,
                  keep_digest=True)
    summ = rep.summary()
    print("serving full range:", {k: summ[k] for k in ("completed", "failed", "p50_s", "p99_s",
                                                       "aggregate_zone_updates_per_s")},
          flush=True)
    assert summ["completed"] + summ["failed"] == len(reqs)
    assert summ["completed"] >= 1 and summ["p99_s"] >= summ["p50_s"] > 0.0
    for r in reqs:
        print(f"  alone: request {r.rid} ({r.dim}D {r.n}, {r.kind})", flush=True)
        if r.error is None:
            tn, tm, dg = S.run_alone(r)
            assert r.steps == tn and r.digest == dg, f"request {r.rid} differs in the mix"
        else:
            with pytest.raises(HgError):
                S.run_alone(r)


@pytest.mark.parametrize("shape", [(4, 0, 0), (2, 2048, 0), (16, 1024, 0), (4, 64, 0)])
def test_batch_fused_and_separate_forms_agree(hiplib, shape):
    """The x-fused batch form (each stage's x sweep with its RK update,
    members alternating between two state sets) ends every member bit for
    bit where the separate form and the member alone end, including a
    member removed while its state sits in the second set."""
    reqs = small_requests(4, seed=17)
    digests = {}
    for fused in (True, False):
        print(f"case batch form fused={fused} shape={shape}", flush=True)
        b = Batch()
        b.configure(*shape)
        b.set_fused(fused)
        engs = []
        for r in reqs:
            prm, g, st = S.setup(r)
            eng = Hydro(**D.desc_fields(prm, g))
            D.load_initial_state(eng, prm, g, st)
            t_end, dt_out, t_out = prm.times()
            b.add(eng, t_end=t_end, t_out=t_out, dt_out=dt_out)
            engs.append((eng, g))
        b.step(1)  # one step of three stages: an odd number of set changes
        assert b.fused() == fused
        eng0, g0 = engs[0]
        b.remove(eng0)  # leaves with its state moved back to its own set
        mid = S.state_digest(eng0, g0)
        while b.step(8):
            pass
        digests[fused] = [mid] + [S.state_digest(e, g) for e, g in engs[1:]]
        b.close()
    assert digests[True] == digests[False]
    for (eng, g), r, dg in zip(engs[1:], reqs[1:], digests[True][1:]):
        assert S.run_alone(r)[2] == dg
