"""Cartesian grid set-up, restating the reference's ``gridset`` and ``metric``.

``gridset`` (src/main/gridset.f90:43-158, crdnt=0, imesh=0) builds the cell
centres ``x1`` and interfaces ``xi1`` by running sums of the spacings, so the
coordinates are not exactly uniform in floating point; the MUSCL limiter
reads them, so the hydro path must receive exactly these arrays.  Each array
covers indices ``s-2 .. e+2`` (ghost width 2), element 0 being index ``s-2``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Axis:
    s: int
    e: int
    x: np.ndarray     # cell centres           x1
    xi: np.ndarray    # right cell interfaces  xi1
    dx: np.ndarray    # centre spacing         dx1
    dxi: np.ndarray   # cell width             dxi1
    idx: np.ndarray   # 1/dx                   idx1
    idxi: np.ndarray  # 1/dxi                  idxi1

    @property
    def n(self) -> int:
        return self.e - self.s + 1


def uniform_axis(xs: float, xe: float, s: int, e: int, gravity: bool = False) -> Axis:
    """One axis of gridset.f90:43-82 (uniform mesh, Cartesian).

    ``gravity`` reproduces the gravswitch/=0 branch (gridset.f90:68-79) for a
    gravity grid equal to the hydro grid: spacings reset to ``dxi``.
    """
    n = e - s + 1
    m = n + 4
    o = s - 2  # index offset: element q holds Fortran index q + o
    dxi = np.full(m, (xe - xs) / float(n))
    dx = np.zeros(m)
    idxi = np.zeros(m)
    idx = np.zeros(m)
    for i in range(s - 1, e + 3):
        q = i - o
        dx[q] = 0.5 * (dxi[q - 1] + dxi[q])
        idxi[q] = 1.0 / dxi[q]
        idx[q] = 1.0 / dx[q]
    # index s-2 is not touched by the reference loop; give it the uniform value
    dx[0] = dxi[0]
    idxi[0] = 1.0 / dxi[0]
    idx[0] = 1.0 / dx[0]
    x = np.zeros(m)
    xi = np.zeros(m)
    x[s - 1 - o] = xs - dxi[s - 1 - o] * 0.5
    x[s - 2 - o] = x[s - 1 - o] - dx[s - 1 - o]
    xi[s - 1 - o] = xs
    xi[s - 2 - o] = xs - dxi[s - 1 - o]
    for i in range(s, e + 3):
        q = i - o
        x[q] = x[q - 1] + dx[q]
        xi[q] = xi[q - 1] + dxi[q]
    if gravity:
        dx = dxi.copy()
        idx = 1.0 / dx
        idxi = 1.0 / dxi
    return Axis(s, e, x, xi, dx, dxi, idx, idxi)


@dataclass
class Grid:
    ax: tuple  # (Axis, Axis, Axis)

    @property
    def shape(self):
        """Interior shape in C order (nz, ny, nx)."""
        return (self.ax[2].n, self.ax[1].n, self.ax[0].n)

    @property
    def padded_shape(self):
        return (self.ax[2].n + 4, self.ax[1].n + 4, self.ax[0].n + 4)

    @property
    def ncells(self) -> int:
        return self.ax[0].n * self.ax[1].n * self.ax[2].n

    @property
    def solve(self):
        return tuple(a.e > a.s for a in self.ax)

    @property
    def dim(self) -> int:
        return sum(self.solve)

    def centres(self, d: int) -> np.ndarray:
        """Interior cell centres of axis d (0-based)."""
        a = self.ax[d]
        return a.x[2:2 + a.n]

    def pad(self, interior: np.ndarray) -> np.ndarray:
        """Interior (nz,ny,nx) array -> padded array with zero ghosts."""
        out = np.zeros(self.padded_shape)
        out[2:-2, 2:-2, 2:-2] = interior
        return out

    @staticmethod
    def interior(padded: np.ndarray) -> np.ndarray:
        return padded[2:-2, 2:-2, 2:-2]


def cartesian_grid(n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), gravity=False) -> Grid:
    """Grid of ``n = (nx, ny, nz)`` cells on [lo, hi] with is=js=ks=1."""
    axes = tuple(uniform_axis(lo[d], hi[d], 1, n[d], gravity) for d in range(3))
    return Grid(axes)
