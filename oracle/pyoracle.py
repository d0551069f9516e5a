"""ctypes wrapper of the CPU oracle (test infrastructure only).

Mirrors :class:`hormone_amd.native.Hydro` so the same host driver can run
either; only tests, the smoke check and the benchmark's CPU-baseline leg use
it, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from hormone_amd.native import HgDesc, HgError, make_desc

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhormone_oracle.so")
_dp = C.POINTER(C.c_double)
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.or_create.argtypes = [C.POINTER(HgDesc), C.POINTER(C.c_int)]
        L.or_create.restype = vp
        L.or_destroy.argtypes = [vp]
        L.or_ufnmax.argtypes = [vp]
        L.or_upload.argtypes = [vp, C.c_int, _dp]
        L.or_download.argtypes = [vp, C.c_int, _dp]
        L.or_pressure.argtypes = [vp]
        L.or_timestep.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_int, C.c_double, _dp, _dp, _dp, _dp]
        L.or_hydro_step.argtypes = [vp, C.c_double, C.c_double]
        L.or_phidamp.argtypes = [vp, C.c_double]
        L.or_last_error.argtypes = [vp] + [C.POINTER(C.c_int)] * 4
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_ghosts.argtypes = [vp, C.c_int, _dp]
        L.or_set_boundary_flux.argtypes = [vp, C.c_int, _dp]
        L.or_gravity.argtypes = [vp, C.c_double, C.c_double, _dp, C.c_double, _dp, _dp,
                                 C.POINTER(C.c_int)]
        L.or_kat_hlldflux.argtypes = [_dp, _dp]
        L.or_kat_get_cf.argtypes = [C.c_double] * 5
        L.or_kat_get_cf.restype = C.c_double
        L.or_kat_eos_p_cs.argtypes = [C.c_double] * 4 + [C.c_int, _dp, _dp]
        L.or_kat_modified_mc.argtypes = [_dp, _dp, _dp]
        L.or_kat_modified_mc.restype = C.c_double
        L.or_kat_minmod.argtypes = [_dp, _dp]
        L.or_kat_minmod.restype = C.c_double
        L.or_kat_get_eint.argtypes = [C.c_int] + [C.c_double] * 8
        L.or_kat_get_eint.restype = C.c_double
        _LIB = L
    return _LIB


def set_threads(n: int):
    lib().or_set_threads(int(n))


class Oracle:
    def __init__(self, **desc_kw):
        self._lib = lib()
        self.desc, self._keep = make_desc(**desc_kw)
        st = C.c_int()
        self.h = self._lib.or_create(C.byref(self.desc), C.byref(st))
        if not self.h:
            raise HgError(st.value, "or_create")
        self._dirty = True
        g = desc_kw["grid"]
        # small grids run faster on one thread than through OpenMP barriers
        if g.ncells < 65536:
            set_threads(1)
        else:
            set_threads(0)

    def close(self):
        if self.h:
            self._lib.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st:
            i, j, k, d = (C.c_int() for _ in range(4))
            self._lib.or_last_error(self.h, C.byref(i), C.byref(j), C.byref(k), C.byref(d))
            raise HgError(st, f"{what} at i={i.value} j={j.value} k={k.value} dir={d.value}",
                          cell=(i.value, j.value, k.value), direction=d.value)

    def ufnmax(self):
        return self._lib.or_ufnmax(self.h)

    def upload(self, field, padded):
        a = np.ascontiguousarray(padded, dtype=np.float64)
        self._check(self._lib.or_upload(self.h, field, a.ctypes.data_as(_dp)), "upload")
        if field < 9 or field >= 16:
            self._dirty = True

    def set_ghosts(self, field, padded):
        a = np.ascontiguousarray(padded, dtype=np.float64)
        self._check(self._lib.or_set_ghosts(self.h, field, a.ctypes.data_as(_dp)), "set_ghosts")

    def set_boundary_flux(self, face, flux):
        a = np.ascontiguousarray(flux, dtype=np.float64)
        assert a.size >= self.ufnmax()
        self._check(self._lib.or_set_boundary_flux(self.h, face, a.ctypes.data_as(_dp)),
                    "set_boundary_flux")

    def download(self, field, padded):
        self._ensure()
        self._check(self._lib.or_download(self.h, field, padded.ctypes.data_as(_dp)), "download")

    def _ensure(self):
        if self._dirty:
            self._check(self._lib.or_pressure(self.h), "pressure")
            self._dirty = False

    def timestep(self, time, t_out, t_end, outstyle, endstyle, fixed_dt=-1.0, grav=False):
        self._ensure()
        dt, ch, cg, dtg = (C.c_double() for _ in range(4))
        self._check(self._lib.or_timestep(self.h, time, t_out, t_end, outstyle, endstyle,
                                          fixed_dt, C.byref(dt), C.byref(ch), C.byref(cg),
                                          C.byref(dtg)), "timestep")
        if grav:
            return dt.value, ch.value, cg.value, dtg.value
        return dt.value, ch.value

    def gravity(self, time, dt, grvtime, cgrav, cgrav_old, dtgrav):
        self._ensure()
        gt, co, dg, ns = C.c_double(grvtime), C.c_double(cgrav_old), C.c_double(dtgrav), \
            C.c_int(0)
        self._check(self._lib.or_gravity(self.h, time, dt, C.byref(gt), cgrav, C.byref(co),
                                         C.byref(dg), C.byref(ns)), "gravity")
        return gt.value, co.value, dg.value, ns.value

    def hydro_step(self, dt, ch):
        self._ensure()
        self._check(self._lib.or_hydro_step(self.h, dt, ch), "hydro_step")

    def phidamp(self, factor):
        self._check(self._lib.or_phidamp(self.h, factor), "phidamp")
