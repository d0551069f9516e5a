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
    for shape in [(4, 0, 0), (1, 512, 1024), (2, 2048, 0), (16, 1024, 0), (8, 0, 2048)]:
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
