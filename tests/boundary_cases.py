"""Problems for the Dirichlet (9) and flux (10) boundary codes, shared by the
CPU tests of the oracle and the GPU tests.

The reference runs neither code on a Cartesian problem (its Dirichlet data
come from the ejecta problem, dirichlet.f90:68-150, and its flux data from
the wind tunnel, fluxboundary.f90:40-88), so the pins are closed forms:

* a uniform flow whose Dirichlet ghosts hold the same state stays uniform
  bit for bit (every face sees identical left and right states, so every
  flux difference is exactly zero) -- and does not with the default ghosts;
* with flux boundaries on both faces of an axis, each conserved total
  changes per step by dt * (F_inner - F_outer) * face area (the interior
  flux differences telescope), whatever the interior does.
"""
import numpy as np

from hormone_amd import driver as D
from hormone_amd import problems as P

GAMMA = 1.4
# p = gamma - 1 makes e_int = p / (gamma - 1) = 1 exactly, and the total
# energy and the e_int the engines derive from it exact too
GM1 = GAMMA - 1.0
V0 = (0.5, 0.25, -0.125)
B0 = (0.5, 0.25, 0.125)


def _bc(axis, code, other=3):
    bc = [other] * 12
    a = "xyz".index(axis)
    for q in (2 * a, 2 * a + 1, 6 + 2 * a, 6 + 2 * a + 1):
        bc[q] = code
    return tuple(bc)


def uniform(cls, axis, code, mag=False, n=48, v=V0, **kw):
    """A uniform state (d = 1, p = gamma-1, velocity v, B = B0 if mag) on a
    1D grid along `axis`, both faces of that axis with boundary code
    `code`."""
    name = ("briowushock_" if mag else "sodshock_") + axis
    prm, g, st = P.make_problem(name, n=n, bc=_bc(axis, code), gamma=GAMMA, **kw)
    st.d[...] = 1.0
    st.p[...] = GM1
    for a in range(3):
        st.v[a][...] = v[a]
        st.b[a][...] = B0[a] if mag else 0.0
    eng = cls(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    return prm, g, eng


def set_uniform_ghosts(eng, g, mag):
    full = lambda v: np.full(g.padded_shape, v)  # noqa: E731
    eng.set_ghosts(D.F_D, full(1.0))
    eng.set_ghosts(D.F_P, full(GM1))
    for a in range(3):
        eng.set_ghosts(D.F_V1 + a, full(V0[a]))
        if mag:
            eng.set_ghosts(D.F_B1 + a, full(B0[a]))


def fields(prm):
    return D.SNAP_FIELDS_MAG if prm.mag_on else D.SNAP_FIELDS


def totals(eng, g, prm):
    """Volume integrals of the conserved variables."""
    from hormone_amd import testnorms as T
    w = T.cell_volumes(g)
    nv = eng.ufnmax()
    out = []
    for q in range(nv):
        a = np.zeros(g.padded_shape)
        eng.download(D.F_U + q, a)
        out.append(float((g.interior(a) * w).sum()))
    return np.array(out)


def face_area(g, axis):
    a = "xyz".index(axis)
    others = [q for q in range(3) if q != a]
    return float(np.prod([g.ax[q].dxi[2:2 + g.ax[q].n].sum() for q in others]))


# fluxes of the two faces (conserved-variable order, padded to 9)
FLUX_IN = np.array([0.01, GM1 + 0.001, 0.002, -0.001, 0.02, 0.003, 0.0, 0.0, 0.0])
FLUX_OUT = np.array([0.004, GM1, 0.0, 0.001, 0.01, 0.0, 0.002, 0.001, 0.0])


def flux_budget(cls, axis, mag, steps=15, n=48):
    """Run `steps` steps with flux boundaries on both faces of `axis` from a
    state at rest; returns (observed change of each conserved total,
    expected change)."""
    prm, g, eng = uniform(cls, axis, 10, mag=mag, n=n, v=(0.0, 0.0, 0.0))
    a = "xyz".index(axis)
    eng.set_boundary_flux(2 * a, FLUX_IN)
    eng.set_boundary_flux(2 * a + 1, FLUX_OUT)
    q0 = totals(eng, g, prm)
    res = D.run(eng, prm, g, max_steps=steps, record_dt=True)
    q1 = totals(eng, g, prm)
    nv = len(q0)
    want = sum(res.dts) * (FLUX_IN[:nv] - FLUX_OUT[:nv]) * face_area(g, axis)
    return q1 - q0, want
