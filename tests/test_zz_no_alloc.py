"""The allocation-free timestep (hg_timestep_noalloc) against hg_timestep.

Caller-owned device arrays hold the primitive fields and the cell widths
(allocated here with hipMalloc from the HIP runtime the library itself is
bound to, so the process holds one runtime); the entry point reduces the
CFL condition in its static scratch and writes dt, ch and a status into a
caller device array.  Two calls in a row on one GPU, the second on a
smaller box, each equal to hg_timestep on a context holding the same state,
bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from hormone_amd import driver as D
from hormone_amd import native
from hormone_amd import problems as P
from hormone_amd.native import Hydro

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceArrays:
    """hipMalloc'd copies of host arrays, freed together."""

    def __init__(self):
        native.lib()  # the library first: its HIP runtime is the one used here
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, n):
        a = np.empty(n, dtype=np.float64)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(a.ctypes.data, p, a.nbytes, D2H) == 0
        return a

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _bind():
    f = native.lib().hg_timestep_noalloc
    f.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_double,
                  C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return f


def _case(dev, name, nsteps, **kw):
    """A context advanced `nsteps` steps, its hg_timestep result at its clock,
    and its primitive state and cell widths copied into device arrays."""
    prm, g, st = P.make_problem(name, **kw)
    eng = Hydro(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    res = D.run(eng, prm, g, max_steps=nsteps)
    t_end, dt_out, t_out = prm.times()
    t_out = max(t_out, res.time + 1e-3 * t_end)
    expect = eng.timestep(res.time, t_out, t_end, prm.outstyle, prm.endstyle)
    fields = (D.F_D, D.F_V1, D.F_V2, D.F_V3, D.F_EINT) + \
        ((D.F_B1, D.F_B2, D.F_B3) if prm.mag_on else ())
    host = D.fetch(eng, g, fields)
    prim = [dev.put(host[f].reshape(-1)) for f in fields]
    dxi = [dev.put(g.ax[a].dxi[2:-2]) for a in range(3)]
    return prm, eng, host, fields, prim, dxi, (res.time, t_out, t_end), expect


def _call(f, desc, prim, dxi, clock, prm, out):
    pp = (C.c_void_p * 8)(*([p.value for p in prim] + [None] * (8 - len(prim))))
    dp = (C.c_void_p * 3)(*[p.value for p in dxi])
    return f(C.byref(desc), pp, dp, clock[0], clock[1], clock[2], prm.outstyle, prm.endstyle,
             out, None)


def _same(a, b):
    return np.float64(a).view(np.uint64) == np.float64(b).view(np.uint64)


def test_noalloc_timestep_equals_hg_timestep(hiplib):
    dev = DeviceArrays()
    f = _bind()
    out = dev.put(np.full(3, -1.0))
    try:
        # the second box is smaller than the first
        for name, nsteps, kw in [("polytrope3d", 3, dict(n=40, steps=10, mag_on=True)),
                                 ("orszagtang_xy", 12, dict(n=48))]:
            print(f"case noalloc {name} {kw}", flush=True)
            prm, eng, _, _, prim, dxi, clock, (dt, ch) = _case(dev, name, nsteps, **kw)
            assert _call(f, eng.desc, prim, dxi, clock, prm, out) == native.HG_OK
            got = dev.get(out, 3)
            assert got[2] == 0.0
            assert _same(got[0], dt) and _same(got[1], ch), (got, dt, ch)
    finally:
        dev.free()


def test_noalloc_timestep_reports_nonpositive_state_and_limits(hiplib):
    dev = DeviceArrays()
    f = _bind()
    out = dev.put(np.zeros(3))
    try:
        print("case noalloc error status and size limit", flush=True)
        prm, eng, host, fields, prim, dxi, clock, _ = _case(dev, "orszagtang_xy", 2, n=32)
        eint = host[D.F_EINT].reshape(-1).copy()
        eint[77] = -1.0  # one cell with e_int < 0
        bad = prim[:4] + [dev.put(eint)] + prim[5:]
        assert _call(f, eng.desc, bad, dxi, clock, prm, out) == native.HG_OK
        assert dev.get(out, 3)[2] == float(native.HG_ERR_NEG_EINT)
        # the next call starts from a cleared status
        assert _call(f, eng.desc, prim, dxi, clock, prm, out) == native.HG_OK
        assert dev.get(out, 3)[2] == 0.0
        # a box beyond the scratch is refused before anything is enqueued
        big, keep = native.make_desc(eng.grid)
        big.ie, big.je, big.ke = big.is_ + 1023, big.js + 1023, big.ks + 64  # 1024*1024*65
        assert _call(f, big, prim, dxi, (0.0, 1.0, 1.0), prm, out) == native.HG_ERR_SCRATCH
    finally:
        dev.free()
