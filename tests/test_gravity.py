"""Cartesian hyperbolic self-gravity: closed-form pins of the oracle.

The reference has no Cartesian branch for hyperbolic gravity
(gravity_hyperbolic.f90:413-417 stops for crdnt=0), so its lap_coeff,
potential boundary and initial state are new here and are pinned by closed
forms instead of reference data:
  * the discrete Laplacian of a quadratic potential is its exact Laplacian;
  * one explicit sub-step obeys psi = dt*c*(lap phi - 4 pi G rho)*damp with
    damp = exp(-(1 - c_old/c + alphagrv*hgcfl)) (sign and factor of the
    source term, gravity_hyperbolic.f90:97-116);
  * an n=1 polytrope in its closed-form potential stays near hydrostatic
    equilibrium (a flipped force or source sign makes it collapse or explode).
"""
import math

import numpy as np

from hormone_amd import driver as D
from hormone_amd import problems as P
from hormone_amd.grid import cartesian_grid
from oracle.pyoracle import Oracle


def quadratic_setup(engine_cls, n=12, rho0=0.7):
    prm = P.Params(n=(n, n, n), lo=(-1.0, -1.0, -1.0), hi=(1.0, 1.0, 1.0), gamma=2.0,
                   bc=(3,) * 12, gravswitch=3, G=1.0, grktype=1, is_test=False)
    g = cartesian_grid(prm.n, prm.lo, prm.hi, gravity=True)
    z = np.zeros(g.shape)
    st = P.State(d=np.full(g.shape, rho0), p=np.full(g.shape, 1.0),
                 v=[z.copy(), z.copy(), z.copy()], b=[z.copy(), z.copy(), z.copy()],
                 phi=z.copy())

    def cen(d):
        return g.ax[d].x
    x, y, zz = np.meshgrid(cen(0), cen(1), cen(2), indexing="ij")
    st.extra["grvphi_padded"] = (x * x + 2.0 * y * y + 3.0 * zz * zz).transpose(2, 1, 0).copy()
    eng = engine_cls(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    return prm, g, eng


def lap_from_one_substep(eng, g, prm, rho0=0.7):
    c, dtg = 2.0, 1e-3
    grvtime, cold, dtg2, nsub = eng.gravity(0.0, dtg, 0.0, c, c, dtg)
    assert nsub == 1 and dtg2 == dtg and cold == c and grvtime == dtg
    psi = D.fetch(eng, g, [D.F_GRVPSI])[D.F_GRVPSI]
    damp = math.exp(-(1.0 - 1.0 + prm.alphagrv * prm.hgcfl))
    return psi / (dtg * c * damp) + 4.0 * math.pi * prm.G * rho0


def test_laplacian_of_quadratic_is_exact():
    prm, g, eng = quadratic_setup(Oracle)
    lap = lap_from_one_substep(eng, g, prm)
    # d2/dx2 + d2/dy2 + d2/dz2 of x^2 + 2y^2 + 3z^2 = 12 on any uniform grid
    assert np.max(np.abs(lap - 12.0)) < 1e-9


def test_polytrope_stays_hydrostatic():
    prm, g, st = P.make_problem("polytrope3d", n=32, steps=20)
    eng = Oracle(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    r = D.run(eng, prm, g)
    assert r.tn == 20 and r.nsub and all(ns >= 1 for ns in r.nsub)
    out = D.fetch(eng, g, [D.F_D, D.F_V1, D.F_V2, D.F_V3, D.F_GRVPHI])
    d = out[D.F_D]
    v2 = out[D.F_V1] ** 2 + out[D.F_V2] ** 2 + out[D.F_V3] ** 2
    vrms = math.sqrt(float((d * v2).sum() / d.sum()))
    vff = math.sqrt(2.0 * 1.0 * 1.0 / 0.3)  # free-fall speed sqrt(2GM/R)
    assert vrms < 0.06 * vff, vrms
    # centre stays the density maximum and the potential minimum
    c = np.unravel_index(np.argmax(d), d.shape)
    assert all(abs(q - 15.5) <= 1.5 for q in c)
    assert out[D.F_GRVPHI].min() < -6.0
