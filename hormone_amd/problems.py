"""Problem set-ups: parameters and initial conditions.

Parameters mirror the reference's namelist files (para/parameters_base plus
para/parameters_<simtype>); initial conditions restate src/init/*.f90 and
the energy initialisation of src/main/initialcondition.f90:133-150.  The
time unit conversion follows checksetup.f90:181-208.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

from .grid import Grid, cartesian_grid

LIMITERS = {"flat": 0, "minmod": 1, "modified_mc": 2}
UNIT_SEC = {"yr": 3600.0 * 24.0 * 365.25, "d": 3600 * 24.0, "hr": 3600.0,
            "ks": 1000.0, "min": 60.0, "s": 1.0, "ms": 1e-3, "ns": 1e-9}


@dataclass
class Params:
    """Settings read by the hot path (defaults: para/parameters_base)."""
    n: tuple = (10, 1, 1)
    lo: tuple = (0.0, 0.0, 0.0)
    hi: tuple = (1.0, 1.0, 1.0)
    outstyle: int = 1
    endstyle: int = 1
    tnlim: int = 0
    tn_out: int = 1
    t_end: float = 0.0       # in dt_unit
    dt_out: float = 1.0      # in dt_unit
    dt_unit: str = "s"
    gamma: float = 1.6666666666666667
    courant: float = 0.9
    rktype: int = 3
    mag_on: bool = False
    flux_limiter: str = "modified_mc"
    alpha9wave: float = 0.1
    fixed_dt: float = -1.0
    # bc1is bc1os bc2is bc2os bc3is bc3os bc1iv bc1ov bc2iv bc2ov bc3iv bc3ov
    bc: tuple = (0,) * 12
    gravswitch: int = 0
    HGfac: float = 5.0
    hgcfl: float = 0.9
    maxtngrv: int = -1
    grktype: int = 2
    alphagrv: float = 0.1
    G: float = 6.67430e-8
    is_test: bool = True
    precision: str = "fp64"   # or "fp32mix"
    compswitch: int = 0       # 2: spn advected species (eos_con)
    spn: int = 0

    @property
    def solve(self):
        return tuple(m > 1 for m in self.n)

    def effective_bc(self):
        """checksetup.f90:126-136: forced codes for unsolved directions."""
        bc = list(self.bc)
        s = self.solve
        if not s[0]:
            for q in (0, 1, 6, 7):
                bc[q] = 2
        if not s[1]:
            for q in (2, 3, 8, 9):
                bc[q] = 3
        if not s[2]:
            for q in (4, 5, 10, 11):
                bc[q] = 3
        return tuple(bc)

    def times(self):
        """(t_end, dt_out, t_out) in seconds, checksetup.f90:201-208."""
        u = UNIT_SEC[self.dt_unit]
        dt_out = self.dt_out * u
        t_end = self.t_end * u
        t_out = min(dt_out, t_end)
        if self.is_test:
            t_out = t_end
        return t_end, dt_out, t_out


# --- parameter files of the reference's Cartesian test problems ----------
# para/parameters_sodshock_{x,y,z}, _briowushock_{x,y,z}, _orszagtang_{..}

def _axis_n(direction: str, n: int):
    return tuple(n if c in direction else 1 for c in "xyz")


def _bc_for(direction: str, code: int):
    bc = [0] * 12
    for d, c in enumerate("xyz"):
        if c in direction:
            for q in (2 * d, 2 * d + 1, 6 + 2 * d, 7 + 2 * d):
                bc[q] = code
    return tuple(bc)


def sodshock_params(direction: str, n: int = 500, **kw) -> Params:
    p = Params(n=_axis_n(direction, n), t_end=200.0, dt_out=200.0, dt_unit="ms",
               gamma=1.4, bc=_bc_for(direction, 3), tnlim=100000)
    return replace(p, **kw)


def briowu_params(direction: str, n: int = 500, **kw) -> Params:
    p = Params(n=_axis_n(direction, n), t_end=200.0, dt_out=200.0, dt_unit="ms",
               gamma=2.0, mag_on=True, bc=_bc_for(direction, 3), tnlim=100000)
    return replace(p, **kw)


def orszagtang_params(plane: str, n: int = 128, **kw) -> Params:
    p = Params(n=_axis_n(plane, n), t_end=500.0, dt_out=500.0, dt_unit="ms",
               mag_on=True, bc=(0,) * 12, tnlim=100000)
    return replace(p, **kw)


def khinstability_params(n: int = 128, **kw) -> Params:
    """para/parameters_KHinstability: 2D hydro shear layer with two advected
    species, periodic, t_end 1000 ms."""
    p = Params(n=(n, n, 1), lo=(-0.5, -0.5, 0.0), hi=(0.5, 0.5, 1.0), t_end=1000.0,
               dt_out=100.0, dt_unit="ms", bc=(0,) * 12, compswitch=2, spn=2, tnlim=100000)
    return replace(p, **kw)


# --- initial conditions ----------------------------------------------------

@dataclass
class State:
    """Interior primitive fields in C order (nz, ny, nx)."""
    d: np.ndarray
    p: np.ndarray
    v: list
    b: list
    phi: np.ndarray
    extra: dict = field(default_factory=dict)


def _zeros(g: Grid):
    return np.zeros(g.shape)


def shocktube(g: Grid, direction: str, kind: str) -> State:
    """src/init/shocktube.f90:14-192 for 'sodshock' and 'briowushock'."""
    dirn = {"x": 1, "y": 2, "z": 3}[direction]
    dl, dr, pl, pr = 1.0, 0.125, 1.0, 0.1
    v1l = v1r = v2l = v2r = v3l = v3r = 0.0
    b1l = b1r = b2l = b2r = b3l = b3r = 0.0
    if kind == "briowushock":
        if dirn == 1:
            b1l, b1r, b2l, b2r, b3l, b3r = 0.75, 0.75, 1.0, -1.0, 0.0, 0.0
        elif dirn == 2:
            b1l, b1r, b2l, b2r, b3l, b3r = 0.0, 0.0, 0.75, 0.75, 1.0, -1.0
        else:
            b1l, b1r, b2l, b2r, b3l, b3r = 1.0, -1.0, 0.0, 0.0, 0.75, 0.75
    ax = g.ax[dirn - 1]
    idx = np.arange(ax.s, ax.e + 1)
    left = idx <= (ax.s + ax.e) // 2
    shape = [1, 1, 1]
    shape[3 - dirn] = ax.n
    left = left.reshape(shape)

    def fill(a, b_):
        return np.broadcast_to(np.where(left, a, b_), g.shape).astype(float).copy()

    return State(d=fill(dl, dr), p=fill(pl, pr),
                 v=[fill(v1l, v1r), fill(v2l, v2r), fill(v3l, v3r)],
                 b=[fill(b1l, b1r), fill(b2l, b2r), fill(b3l, b3r)],
                 phi=_zeros(g))


def orszagtang(g: Grid, gamma: float) -> State:
    """src/init/orszagtang.f90:14-83 (xy, xz or yz plane)."""
    pi = math.acos(-1.0)
    n = [a.n for a in g.ax]
    if n[2] == 1:
        which = 1
    if n[1] == 1:
        which = 2
    if n[0] == 1:
        which = 3
    nz, ny, nx = g.shape
    d = np.full(g.shape, gamma * gamma / (4.0 * pi))
    p = np.full(g.shape, gamma / (4.0 * pi))
    s4 = math.sqrt(4.0 * pi)
    v = [_zeros(g), _zeros(g), _zeros(g)]
    b = [_zeros(g), _zeros(g), _zeros(g)]
    # (first, second) in-plane axes and the Fortran loops' coordinates
    if which == 1:
        ax1, ax2 = 0, 1
    elif which == 2:
        ax1, ax2 = 0, 2
    else:
        ax1, ax2 = 1, 2
    c1 = g.centres(ax1)
    c2 = g.centres(ax2)
    s2a = np.array([-math.sin(2.0 * pi * t) for t in c2])   # -sin(2 pi x_2nd)
    s1a = np.array([math.sin(2.0 * pi * t) for t in c1])    # sin(2 pi x_1st)
    s4a = np.array([math.sin(4.0 * pi * t) / s4 for t in c1])
    bs2 = np.array([-math.sin(2.0 * pi * t) / s4 for t in c2])

    def along(arr, axis):
        shape = [1, 1, 1]
        shape[2 - axis] = arr.size
        return np.broadcast_to(arr.reshape(shape), g.shape).copy()

    v[ax1] = along(s2a, ax2)
    v[ax2] = along(s1a, ax1)
    b[ax1] = along(bs2, ax2)
    b[ax2] = along(s4a, ax1)
    return State(d=d, p=p, v=v, b=b, phi=_zeros(g))


def khinstability(g: Grid) -> State:
    """src/init/KHtests.f90:14-56: tanh-like shear layer |y| = 1/4 of width
    0.025 (exp profiles), species 1 outside and 2 inside the layer,
    vy = 0.01 sin(4 pi x), p = 2.5.  exp and sin are the C library's, as the
    reference's."""
    pi = math.acos(-1.0)
    rho1, rho2 = 1.0, 2.0
    rhom = 0.5 * (rho1 - rho2)
    vv1, vv2 = 0.5, -0.5
    vvm = 0.5 * (vv1 - vv2)
    ptb, width = 1e-2, 0.025
    ys = g.centres(1)
    v1y, dy, s1y = [], [], []
    for y in ys:
        ay = abs(float(y))
        if ay > 0.25:
            ex = math.exp((-ay + 0.25) / width)
            v1y.append(vv1 - vvm * ex)
            dy.append(rho1 - rhom * ex)
            s1y.append(1.0)
        else:
            ex = math.exp((ay - 0.25) / width)
            v1y.append(vv2 + vvm * ex)
            dy.append(rho2 + rhom * ex)
            s1y.append(0.0)
    sx = np.array([ptb * math.sin(4.0 * pi * float(x)) for x in g.centres(0)])

    def along_y(a):
        return np.broadcast_to(np.asarray(a)[None, :, None], g.shape).copy()

    s1 = along_y(s1y)
    z = np.zeros(g.shape)
    return State(d=along_y(dy), p=np.full(g.shape, 2.5),
                 v=[along_y(v1y), np.broadcast_to(sx[None, None, :], g.shape).copy(), z.copy()],
                 b=[z.copy(), z.copy(), z.copy()], phi=z.copy(),
                 extra={"spc": [s1, 1.0 - s1], "species": ["scalar1", "scalar2"]})


def initial_energy(st: State, gamma: float):
    """initialcondition.f90:133-150 (eostype 0): eint = p/(gamma-1),
    e = eint + d*0.5*|v|^2 + 0.5*|B|^2, evaluated left to right."""
    v1, v2, v3 = st.v
    b1, b2, b3 = st.b
    eint = st.p / (gamma - 1.0)
    e = (eint + (st.d * 0.5) * ((v1 * v1 + v2 * v2) + v3 * v3)) \
        + 0.5 * ((b1 * b1 + b2 * b2) + b3 * b3)
    return eint, e


def conserve(st: State, e: np.ndarray, mag_on: bool, dim: int):
    """conserve.f90:13-49: the conserved variables in equation order."""
    u = [st.d.copy(), st.d * st.v[0], st.d * st.v[1], st.d * st.v[2], e.copy()]
    if mag_on:
        u += [st.b[0].copy(), st.b[1].copy(), st.b[2].copy()]
        if dim > 1:
            u.append(st.phi.copy())
    return u


# --- seeded counter-based noise (same value for a cell whatever the layout) --

_M64 = (1 << 64) - 1


def hash_uniform(seed: int, i, j, k) -> np.ndarray:
    """Uniform numbers in [-1, 1) from a splitmix64-style hash of
    (seed, i, j, k), with i, j, k the Fortran cell indices."""
    with np.errstate(over="ignore"):
        x = (np.uint64(seed & _M64) * np.uint64(0x9E3779B97F4A7C15)
             + np.asarray(i, dtype=np.uint64) * np.uint64(0xBF58476D1CE4E5B9)
             + np.asarray(j, dtype=np.uint64) * np.uint64(0x94D049BB133111EB)
             + np.asarray(k, dtype=np.uint64) * np.uint64(0xD6E8FEB86659FD93))
        x = x ^ (x >> np.uint64(30))
        x = x * np.uint64(0xBF58476D1CE4E5B9)
        x = x ^ (x >> np.uint64(27))
        x = x * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return (x >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def _cell_indices(g: Grid, pad: int = 0):
    """Fortran indices (i, j, k) broadcast to the (nz, ny, nx) layout."""
    ii = np.arange(g.ax[0].s - pad, g.ax[0].e + 1 + pad)
    jj = np.arange(g.ax[1].s - pad, g.ax[1].e + 1 + pad)
    kk = np.arange(g.ax[2].s - pad, g.ax[2].e + 1 + pad)
    return np.meshgrid(ii, jj, kk, indexing="ij")


# --- n=1 polytrope with hyperbolic self-gravity (new Cartesian set-up) -----

@dataclass
class Polytrope:
    M: float = 1.0
    R: float = 0.3
    G: float = 1.0
    rho_bg: float = 1e-6
    noise: float = 1e-3
    seed: int = 12345

    @property
    def rho_c(self):
        return math.pi * self.M / (4.0 * self.R ** 3)

    @property
    def K(self):
        return 2.0 * self.G * self.R ** 2 / math.pi

    def rho(self, r):
        xi = math.pi * r / self.R
        with np.errstate(invalid="ignore", divide="ignore"):
            inside = np.where(xi > 0, np.sin(xi) / np.where(xi > 0, xi, 1.0), 1.0)
        return np.where(r < self.R, self.rho_c * inside, 0.0)

    def potential(self, r):
        """Closed-form potential: -GM/R - 2K rho inside, -GM/r outside."""
        inside = -self.G * self.M / self.R - 2.0 * self.K * self.rho(r)
        with np.errstate(divide="ignore"):
            outside = -self.G * self.M / np.where(r > 0, r, 1.0)
        return np.where(r < self.R, inside, outside)


def polytrope_params(n: int = 256, mag_on: bool = False, steps: int = 200, **kw) -> Params:
    p = Params(n=(n, n, n), lo=(-1.0, -1.0, -1.0), hi=(1.0, 1.0, 1.0), gamma=2.0,
               mag_on=mag_on, bc=(3,) * 12, gravswitch=3, G=1.0, outstyle=2, endstyle=2,
               tnlim=steps, tn_out=steps, t_end=1e30, is_test=False)
    return replace(p, **kw)


def _radius(g: Grid, pad: int = 0):
    def cen(d):
        a = g.ax[d]
        return a.x[2 - pad:2 + a.n + pad]
    x, y, z = np.meshgrid(cen(0), cen(1), cen(2), indexing="ij")
    return np.sqrt(x * x + y * y + z * z).transpose(2, 1, 0)


def polytrope(g: Grid, pol: Polytrope = Polytrope()) -> State:
    """n=1 polytrope (gamma=2) of mass M, radius R in a uniform background,
    density perturbed by seeded uniform noise; potential in closed form."""
    r = _radius(g)
    rp = pol.rho(r)
    rho0 = np.maximum(rp, pol.rho_bg)
    ii, jj, kk = _cell_indices(g)
    u = hash_uniform(pol.seed, ii, jj, kk).transpose(2, 1, 0)
    d = rho0 * (1.0 + pol.noise * u)
    # hydrostatic pressure: K rho^2 in the star plus the background's own
    # hydrostatic profile rho_bg*G*M/max(r,R) (continuous at r = R)
    p = pol.K * rp * rp + pol.rho_bg * pol.G * pol.M / np.maximum(r, pol.R)
    z = np.zeros(g.shape)
    rp = _radius(g, pad=2)
    phi_pad = pol.potential(rp)
    return State(d=d, p=p, v=[z.copy(), z.copy(), z.copy()], b=[z.copy(), z.copy(), z.copy()],
                 phi=z.copy(), extra={"grvphi_padded": phi_pad})


# --- 3D MHD Kelvin-Helmholtz shear layer (new set-up) -----------------------

def kh_params(n: int = 384, steps: int = 100, **kw) -> Params:
    p = Params(n=(n, n, n), lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), gamma=5.0 / 3.0,
               mag_on=True, bc=(0,) * 12, outstyle=2, endstyle=2, tnlim=steps, tn_out=steps,
               t_end=1e30, is_test=False)
    return replace(p, **kw)


def kelvin_helmholtz(g: Grid, seed: int = 2024, amp: float = 0.01) -> State:
    """rho = 1, vx = +0.5 for |y - 1/2| < 1/4, rho = 2, vx = -0.5 outside;
    p = 2.5, B = (0.1, 0, 0); vy = amp * U(-1,1) * sin(4 pi x), U from the
    seeded counter hash of the cell indices."""
    x = g.centres(0)[None, None, :]
    y = g.centres(1)[None, :, None]
    inside = np.abs(y - 0.5) < 0.25
    d = np.broadcast_to(np.where(inside, 1.0, 2.0), g.shape).copy()
    vx = np.broadcast_to(np.where(inside, 0.5, -0.5), g.shape).copy()
    ii, jj, kk = _cell_indices(g)
    u = hash_uniform(seed, ii, jj, kk).transpose(2, 1, 0)
    vy = amp * u * np.sin(4.0 * math.pi * np.broadcast_to(x, g.shape))
    z = np.zeros(g.shape)
    return State(d=d, p=np.full(g.shape, 2.5), v=[vx, vy, z.copy()],
                 b=[np.full(g.shape, 0.1), z.copy(), z.copy()], phi=z.copy())


def make_problem(name: str, **kw):
    """(Params, Grid, State) for a reference test problem name."""
    if name.startswith("sodshock_"):
        prm = sodshock_params(name[-1], **kw)
        g = cartesian_grid(prm.n, prm.lo, prm.hi)
        return prm, g, shocktube(g, name[-1], "sodshock")
    if name.startswith("briowushock_"):
        prm = briowu_params(name[-1], **kw)
        g = cartesian_grid(prm.n, prm.lo, prm.hi)
        return prm, g, shocktube(g, name[-1], "briowushock")
    if name.startswith("orszagtang_"):
        prm = orszagtang_params(name.split("_")[1], **kw)
        g = cartesian_grid(prm.n, prm.lo, prm.hi)
        return prm, g, orszagtang(g, prm.gamma)
    if name == "KHinstability":
        prm = khinstability_params(**kw)
        g = cartesian_grid(prm.n, prm.lo, prm.hi)
        return prm, g, khinstability(g)
    if name == "kh3d":
        seed = kw.pop("seed", 2024)
        prm = kh_params(**kw)
        g = cartesian_grid(prm.n, prm.lo, prm.hi)
        return prm, g, kelvin_helmholtz(g, seed)
    if name == "polytrope3d":
        pol = kw.pop("polytrope", Polytrope())
        prm = polytrope_params(**kw)
        g = cartesian_grid(prm.n, prm.lo, prm.hi, gravity=True)
        return prm, g, polytrope(g, pol)
    raise ValueError("unknown problem " + name)
