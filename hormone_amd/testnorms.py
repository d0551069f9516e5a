"""The reference's test criterion for a run against a golden snapshot
(tests.f90:48-298): per-variable max, "L1" and discontinuity-weighted L2
norm errors, and the pass rule maxval(L2) < test_tol.

Host-side mirror of `test`, `max_norm_error`, `L1_norm_error`,
`L2_norm_error` and `ratio`; arrays are (nz, ny, nx) as in the rest of the
host code, weights are the cell volumes dvol = dxi1*dxi2*dxi3 (Cartesian,
metric.f90:44-47).

Differences from the reference, both in how fields reach the norms, not in
the norms:
* species: the reference stores species n at column 10+n but labels column
  11+n with its name (tests.f90:75-79,103-107), so with radiation off the
  first species is skipped and the last label reads an unset column; here
  species n is compared as itself.
* cs for the velocity scale comes from the golden state (the reference
  calls timestep after readbin, tests.f90:131, so its cs is the golden's
  too), via eos_p_cs for eostype 0: cs = sqrt(gamma*p/d).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.ndimage import maximum_filter, minimum_filter

EPS = float(np.finfo(np.float64).eps)  # epsilon(1d0)
DISCO_RANGE = 3  # tests.f90:258

# &testcon of each problem's parameter file (para/parameters_*); Mach_tol
# from parameters_base
TEST_TOL = {"sodshock": 1e-15, "briowushock": 1e-15, "orszagtang": 7e-15,
            "KHinstability": 1e-15}
MACH_TOL = 1e-5


def tolerance_for(problem: str) -> float:
    """test_tol of the problem's parameter file."""
    for k, v in TEST_TOL.items():
        if problem.startswith(k):
            return v
    raise KeyError(problem)


def max_norm_error(var, var0, scale) -> float:
    """tests.f90:198-213."""
    pos = var.max() * var.min()
    with np.errstate(divide="ignore", invalid="ignore"):
        if pos > 0.0:
            return float(np.abs(var / var0 - 1.0).max())
        return float((np.abs(var - var0) / np.maximum(np.abs(var0), scale)).max())


def l1_norm_error(var, var0, scale, w) -> float:
    """tests.f90:215-241 (a square-root form despite the name)."""
    pos = var.max() * var.min()
    with np.errstate(divide="ignore", invalid="ignore"):
        if pos > 0.0:
            errsum = (np.abs(var / var0 - 1.0) * w).sum()
        else:
            errsum = (np.abs(var - var0) / np.maximum(np.abs(var0), scale) * w).sum()
    return math.sqrt(errsum) / math.sqrt(w.sum())


def ratio(x, y, floor):
    """tests.f90:291-295: max(|x|/max(|y|,floor), |y|/max(|x|,floor), 1)."""
    return np.maximum(np.maximum(np.abs(x) / np.maximum(np.abs(y), floor),
                                 np.abs(y) / np.maximum(np.abs(x), floor)), 1.0)


def l2_norm_error(var, var0, scale, w) -> float:
    """tests.f90:243-289.  Each cell's error is normalised by the largest
    |var0| within DISCO_RANGE cells (the window is clipped at the domain
    edge) and weighed down by exp(1 - jump), jump being the largest ratio
    between the cell's value and any value in the window.  The window
    maximum of ratio(x, base, floor) is evaluated from the window's extreme
    |x| -- division is monotone, so the result is the same."""
    pos = var.max() * var.min()
    size = 2 * DISCO_RANGE + 1
    a0 = np.abs(var0)
    wmax = maximum_filter(a0, size=size, mode="nearest")  # maxval(abs(var0(window)))
    wmin = minimum_filter(a0, size=size, mode="nearest")
    base = var0
    denom = wmax
    floor = np.zeros_like(var0)
    if pos <= 0.0:
        base = np.maximum(np.abs(var0), scale)
        denom = np.maximum(denom, scale)
        floor = scale
    with np.errstate(divide="ignore", invalid="ignore"):
        ab = np.abs(base)
        jump = np.maximum(np.maximum(wmax / np.maximum(ab, floor),
                                     ab / np.maximum(wmin, floor)), 1.0)
        relerr = (var - var0) / denom * np.exp(1.0 - jump)
    return math.sqrt((relerr ** 2 * w).sum()) / math.sqrt(w.sum())


def cell_volumes(g) -> np.ndarray:
    """dvol = dxi1*dxi2*dxi3 of the interior cells (metric.f90:44-47), (nz, ny, nx)."""
    d1, d2, d3 = (g.ax[a].dxi[2:2 + g.ax[a].n] for a in range(3))
    return (d1[None, None, :] * d2[None, :, None]) * d3[:, None, None]


def reference_criterion(state: dict, golden: dict, gamma: float, w, tol: float,
                        mag_on: bool, dim: int, species=(), mach_tol: float = MACH_TOL):
    """The reference's `test` (tests.f90:48-178) for hydro/MHD fields.

    state, golden: dicts of (nz, ny, nx) arrays with keys d e v1 v2 v3
    [b1 b2 b3 phi] [species...].  w: cell volumes.  Returns (passed,
    {norm: {variable: error}}, identical): passed is maxval(L2) < tol and
    identical is maxval(L2) <= 0, the reference's "test: passed" and
    "Identical!" (tests.f90:166-176)."""
    names = ["d", "e", "v1", "v2", "v3"]
    if mag_on:
        names += ["b1", "b2", "b3"]
        if dim >= 2:
            names.append("phi")
    names += list(species)
    g = golden
    vmag = np.sqrt(g["v1"] ** 2 + g["v2"] ** 2 + g["v3"] ** 2)
    bmag = (np.sqrt(g["b1"] ** 2 + g["b2"] ** 2 + g["b3"] ** 2) if mag_on
            else np.zeros_like(vmag))
    eint = g["e"] - 0.5 * g["d"] * (g["v1"] ** 2 + g["v2"] ** 2 + g["v3"] ** 2)
    if mag_on:
        eint = eint - 0.5 * (g["b1"] ** 2 + g["b2"] ** 2 + g["b3"] ** 2)
    cs = np.sqrt(gamma * ((gamma - 1.0) * eint) / g["d"])
    out = {"max": {}, "L1": {}, "L2": {}}
    for nm in names:
        if nm in ("v1", "v2", "v3"):
            scale = np.maximum(vmag, cs * mach_tol)
        elif nm in ("b1", "b2", "b3"):
            scale = bmag
        else:
            scale = np.abs(g[nm])
        scale = np.maximum(scale, EPS)
        var, var0 = np.asarray(state[nm], np.float64), np.asarray(g[nm], np.float64)
        out["max"][nm] = max_norm_error(var, var0, scale)
        out["L1"][nm] = l1_norm_error(var, var0, scale, w)
        out["L2"][nm] = l2_norm_error(var, var0, scale, w)
    worst = max(out["L2"].values())
    return bool(worst < tol), out, bool(worst <= 0.0)
