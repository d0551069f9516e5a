"""The reference's test criterion (tests.f90:48-298) as mirrored in
hormone_amd/testnorms.py: pinned against a literal transcription of the
Fortran loops on small arrays, closed forms, and the oracle's run of a
reference golden."""
import math

import numpy as np
import pytest

from hormone_amd import testnorms as T
from hormone_amd.grid import cartesian_grid


def _l2_literal(var, var0, scale, w):
    """tests.f90:243-289 statement by statement (window clipped at the edges,
    ratio over every cell of the window)."""
    pos = var.max() * var.min()
    nz, ny, nx = var.shape
    r = T.DISCO_RANGE
    relerr = np.zeros_like(var)
    for k in range(nz):
        kl, ku = min(r, k), min(r, nz - 1 - k)
        for j in range(ny):
            jl, ju = min(r, j), min(r, ny - 1 - j)
            for i in range(nx):
                il, iu = min(r, i), min(r, nx - 1 - i)
                win = var0[k - kl:k + ku + 1, j - jl:j + ju + 1, i - il:i + iu + 1]
                base = var0[k, j, i]
                denom = np.abs(win).max()
                floor = 0.0
                if pos <= 0.0:
                    base = max(abs(base), scale[k, j, i])
                    denom = max(denom, scale[k, j, i])
                    floor = scale[k, j, i]
                with np.errstate(divide="ignore"):
                    jump = T.ratio(win, base, floor).max()
                relerr[k, j, i] = (var[k, j, i] - var0[k, j, i]) / denom * math.exp(1.0 - jump)
    return math.sqrt((relerr ** 2 * w).sum()) / math.sqrt(w.sum())


def _fields(rng, shape, positive):
    x = rng.uniform(0.5, 2.0, shape) if positive else rng.uniform(-1, 1, shape)
    if not positive:
        x[x > 0.6] = 0.0  # zeros, as velocities and fields have
    x[..., shape[-1] // 2:] *= 3.0  # a jump
    return x


@pytest.mark.parametrize("shape", [(1, 1, 40), (1, 9, 11), (5, 6, 7)])
@pytest.mark.parametrize("positive", [True, False])
def test_l2_matches_literal_loops(shape, positive):
    rng = np.random.default_rng(3)
    var0 = _fields(rng, shape, positive)
    var = var0 * (1 + 1e-6 * rng.standard_normal(shape)) + 1e-9 * (var0 == 0)
    scale = np.maximum(np.abs(var0), T.EPS)
    w = rng.uniform(0.5, 1.5, shape)
    got = T.l2_norm_error(var, var0, scale, w)
    want = _l2_literal(var, var0, scale, w)
    assert got == pytest.approx(want, rel=1e-14, abs=0)


def test_closed_forms():
    g = cartesian_grid((32, 1, 1))
    w = T.cell_volumes(g)
    x = g.centres(0)[None, None, :]
    var0 = 1.0 + 0.1 * np.sin(2 * np.pi * x)
    scale = np.maximum(np.abs(var0), T.EPS)
    assert T.max_norm_error(var0, var0, scale) == 0.0
    assert T.l1_norm_error(var0, var0, scale, w) == 0.0
    assert T.l2_norm_error(var0, var0, scale, w) == 0.0
    # one cell off by delta: max = delta/v0, "L1" = sqrt(delta/v0 * w/W)
    var = var0.copy()
    c = 16
    var[0, 0, c] *= 1.0 + 1e-8
    rel = abs(var[0, 0, c] / var0[0, 0, c] - 1.0)
    assert T.max_norm_error(var, var0, scale) == pytest.approx(rel, rel=1e-12)
    assert T.l1_norm_error(var, var0, scale, w) == pytest.approx(math.sqrt(rel / 32), rel=1e-12)
    # L2 in a smooth region: denominator = window max, jump = window max/min ratio
    win = var0[0, 0, c - 3:c + 4]
    jump = max(win.max() / var0[0, 0, c], var0[0, 0, c] / win.min(), 1.0)
    want = abs(var[0, 0, c] - var0[0, 0, c]) / win.max() * math.exp(1 - jump) / math.sqrt(32)
    assert T.l2_norm_error(var, var0, scale, w) == pytest.approx(want, rel=1e-12)


def test_l2_weighs_down_errors_next_to_a_jump():
    g = cartesian_grid((64, 1, 1))
    w = T.cell_volumes(g)
    var0 = np.where(np.arange(64) < 32, 1.0, 10.0)[None, None, :]
    scale = np.abs(var0)
    near, far = var0.copy(), var0.copy()
    near[0, 0, 30] += 1e-6
    far[0, 0, 10] += 1e-6
    e_near = T.l2_norm_error(near, var0, scale, w)
    e_far = T.l2_norm_error(far, var0, scale, w)
    assert e_near == pytest.approx(e_far / 10.0 * math.exp(1.0 - 10.0), rel=1e-12)


def test_oracle_run_passes_the_reference_criterion():
    """The oracle's Sod run against the reference's golden: the reference's
    own criterion reports 'Identical!' (every L2 error is zero)."""
    from oracle.pyoracle import Oracle
    from tests.helpers import GOLDEN, run_problem
    res, state, _ = run_problem(Oracle, "sodshock_x")
    z = np.load(f"{GOLDEN}/sodshock_x.npz")
    shape = state["d"].shape
    golden = {k: z[k].reshape(shape) for k in ("d", "v1", "v2", "v3", "e")}
    g = cartesian_grid((shape[-1], 1, 1))
    passed, errs, identical = T.reference_criterion(
        state, golden, 1.4, T.cell_volumes(g), T.tolerance_for("sodshock_x"), False, 1)
    assert passed and identical, errs
    # and fails a run off by a relative 1e-13 in one cell of the smooth left
    # state: L2 = 1e-13/sqrt(500) = 4.5e-15 > test_tol = 1e-15
    bad = dict(state)
    bad["d"] = state["d"].copy()
    bad["d"].reshape(-1)[50] *= 1 + 1e-13
    passed, errs, identical = T.reference_criterion(
        bad, golden, 1.4, T.cell_volumes(g), T.tolerance_for("sodshock_x"), False, 1)
    assert not passed and not identical
    assert errs["L2"]["d"] == pytest.approx(1e-13 / math.sqrt(500), rel=0.05)
