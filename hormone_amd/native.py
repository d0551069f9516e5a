"""ctypes binding of the C ABI declared in include/hormone_hip.h.

The HIP library ``libhormone_hip.so`` is built in-tree next to this file by
``__graft_entry__.build()``; there is no fallback: constructing
:class:`Hydro` without it raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# HORMONE_HIP_LIB names another build of the same library (a variant built
# with different compile-time options); the default is the in-tree build.
LIB_PATH = os.environ.get("HORMONE_HIP_LIB") or os.path.join(HERE, "libhormone_hip.so")

HG_OK, HG_ERR_ARG, HG_ERR_UNSUPPORTED, HG_ERR_DEVICE = 0, 1, 2, 3
HG_ERR_FACE_EOS, HG_ERR_NEG_EINT, HG_ERR_NOMEM, HG_ERR_STATE = 4, 5, 6, 7
HG_ERR_SCRATCH = 8
PREC_FP64, PREC_FP32MIX = 0, 1

_dp = C.POINTER(C.c_double)


class HgDesc(C.Structure):
    _fields_ = [
        ("is_", C.c_int32), ("ie", C.c_int32), ("js", C.c_int32),
        ("je", C.c_int32), ("ks", C.c_int32), ("ke", C.c_int32),
        ("x1", _dp), ("xi1", _dp), ("dx1", _dp), ("dxi1", _dp), ("idx1", _dp), ("idxi1", _dp),
        ("x2", _dp), ("xi2", _dp), ("dx2", _dp), ("dxi2", _dp), ("idx2", _dp), ("idxi2", _dp),
        ("x3", _dp), ("xi3", _dp), ("dx3", _dp), ("dxi3", _dp), ("idx3", _dp), ("idxi3", _dp),
        ("crdnt", C.c_int32), ("eostype", C.c_int32), ("compswitch", C.c_int32),
        ("mag_on", C.c_int32), ("flux_limiter", C.c_int32), ("rktype", C.c_int32),
        ("bc", C.c_int32 * 12),
        ("gamma", C.c_double), ("courant", C.c_double), ("alpha9wave", C.c_double),
        ("gravswitch", C.c_int32), ("grktype", C.c_int32),
        ("HGfac", C.c_double), ("hgcfl", C.c_double), ("alphagrv", C.c_double),
        ("G", C.c_double), ("maxtngrv", C.c_int32), ("precision", C.c_int32),
        ("device", C.c_int32), ("spn", C.c_int32),
    ]


class HgError(RuntimeError):
    def __init__(self, code, where="", cell=None, direction=None):
        self.code = code
        self.cell = cell            # (i, j, k) in the reference's Fortran indices
        self.direction = direction  # 1..3 for a face error, 0 for a cell error
        self.tn = None
        self.time = None
        super().__init__(f"hormone_hip status {code}{(': ' + where) if where else ''}")


def make_desc(grid, mag_on=0, flux_limiter=2, rktype=3, bc=(0,) * 12,
              gamma=5.0 / 3.0, courant=0.9, alpha9wave=0.1, gravswitch=0,
              grktype=2, HGfac=5.0, hgcfl=0.9, alphagrv=0.1, G=6.67430e-8,
              maxtngrv=-1, precision=PREC_FP64, device=0, compswitch=0, spn=0):
    """Build an HgDesc; returns (desc, keepalive) -- keep both referenced."""
    d = HgDesc()
    keep = []
    names = ("x", "xi", "dx", "dxi", "idx", "idxi")
    for a in range(3):
        ax = grid.ax[a]
        setattr(d, ("is_", "js", "ks")[a], ax.s)
        setattr(d, ("ie", "je", "ke")[a], ax.e)
        for nm in names:
            arr = np.ascontiguousarray(getattr(ax, nm), dtype=np.float64)
            keep.append(arr)
            setattr(d, f"{nm}{a + 1}", arr.ctypes.data_as(_dp))
    d.crdnt = 0
    d.eostype = 0
    d.compswitch = int(compswitch)
    d.spn = int(spn)
    d.mag_on = int(mag_on)
    d.flux_limiter = int(flux_limiter)
    d.rktype = int(rktype)
    for q in range(12):
        d.bc[q] = int(bc[q])
    d.gamma = gamma
    d.courant = courant
    d.alpha9wave = alpha9wave
    d.gravswitch = int(gravswitch)
    d.grktype = int(grktype)
    d.HGfac, d.hgcfl, d.alphagrv, d.G = HGfac, hgcfl, alphagrv, G
    d.maxtngrv = int(maxtngrv)
    d.precision = int(precision)
    d.device = int(device)
    return d, keep


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    lib.hg_version.restype = C.c_char_p
    lib.hg_create.argtypes = [C.POINTER(HgDesc), C.POINTER(vp)]
    lib.hg_destroy.argtypes = [vp]
    lib.hg_ufnmax.argtypes = [vp]
    lib.hg_upload.argtypes = [vp, C.c_int32, _dp]
    lib.hg_download.argtypes = [vp, C.c_int32, _dp]
    lib.hg_timestep.argtypes = [vp, C.c_double, C.c_double, C.c_double,
                                C.c_int32, C.c_int32, C.c_double, _dp, _dp,
                                _dp, _dp]
    lib.hg_hydro_step.argtypes = [vp, C.c_double, C.c_double]
    lib.hg_phidamp.argtypes = [vp, C.c_double]
    lib.hg_last_error.argtypes = [vp] + [C.POINTER(C.c_int32)] * 4
    lib.hg_run.argtypes = [vp, _dp, C.POINTER(C.c_int64), _dp, C.c_double, C.c_double,
                           C.c_int32, C.c_int32, C.c_int64, C.c_int64, C.c_int64,
                           C.c_double, C.POINTER(C.c_int32)]
    lib.hg_run_gravity.argtypes = [vp, _dp, C.POINTER(C.c_int64), _dp, C.c_double,
                                   C.c_double, C.c_int32, C.c_int32, C.c_int64, C.c_int64,
                                   C.c_int64, C.c_double, _dp, _dp, C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int32)]
    lib.hg_profile.argtypes = [vp, C.c_int32]
    lib.hg_profile_read.argtypes = [vp, C.c_int32, _dp, C.POINTER(C.c_int64), _dp]
    lib.hg_kat.argtypes = [C.c_int32, C.c_int32, _dp, _dp]
    lib.hg_gravity.argtypes = [vp, C.c_double, C.c_double, _dp, C.c_double, _dp, _dp,
                               C.POINTER(C.c_int32)]
    lib.hg_set_launch.argtypes = [vp, C.c_int32]
    lib.hg_time_sweep.argtypes = [vp, C.c_int32, _dp]
    lib.hg_set_fused.argtypes = [vp, C.c_int32]
    lib.hg_set_ghosts.argtypes = [vp, C.c_int32, _dp]
    lib.hg_set_boundary_flux.argtypes = [vp, C.c_int32, _dp]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def exported_symbols():
    """Entry points declared in include/hormone_hip.h."""
    return ["hg_version", "hg_create", "hg_destroy", "hg_ufnmax", "hg_upload",
            "hg_download", "hg_timestep", "hg_hydro_step", "hg_phidamp",
            "hg_last_error", "hg_run", "hg_run_gravity", "hg_profile", "hg_profile_read", "hg_kat",
            "hg_gravity", "hg_set_launch", "hg_time_sweep", "hg_set_fused",
            "hg_set_ghosts", "hg_set_boundary_flux"]


STOP_END, STOP_OUTPUT, STOP_MAXSTEPS, STOP_ERROR = 1, 2, 3, 4
PROF_SWEEP, PROF_RK, PROF_CFL, PROF_GRAV = 0, 1, 2, 3
PROF_ONLY = 0x100


class Hydro:
    """One device-resident simulation context (hg_ctx)."""

    def __init__(self, **desc_kw):
        self._lib = lib()
        self.desc, self._keep = make_desc(**desc_kw)
        self.grid = desc_kw["grid"]
        h = C.c_void_p()
        st = self._lib.hg_create(C.byref(self.desc), C.byref(h))
        if st != HG_OK:
            raise HgError(st, "hg_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self._lib.hg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what, **extra):
        if st != HG_OK:
            i, j, k, d = (C.c_int32() for _ in range(4))
            self._lib.hg_last_error(self.h, C.byref(i), C.byref(j), C.byref(k), C.byref(d))
            err = HgError(st, f"{what} at i={i.value} j={j.value} k={k.value} dir={d.value}",
                          cell=(i.value, j.value, k.value), direction=d.value)
            for key, val in extra.items():
                setattr(err, key, val)
            raise err

    def ufnmax(self):
        return self._lib.hg_ufnmax(self.h)

    def upload(self, field, padded):
        a = np.ascontiguousarray(padded, dtype=np.float64)
        self._check(self._lib.hg_upload(self.h, field, a.ctypes.data_as(_dp)), "hg_upload")

    def set_ghosts(self, field, padded):
        """Ghost values of `field` at the faces with codes 4, 5, 9, 10
        (hg_set_ghosts); `padded` is shaped like an upload."""
        a = np.ascontiguousarray(padded, dtype=np.float64)
        self._check(self._lib.hg_set_ghosts(self.h, field, a.ctypes.data_as(_dp)),
                    "hg_set_ghosts")

    def set_boundary_flux(self, face, flux):
        """Flux of flux-boundary face `face` (0..5 = x1i x1o x2i x2o x3i x3o)."""
        a = np.ascontiguousarray(flux, dtype=np.float64)
        assert a.size >= self.ufnmax()
        self._check(self._lib.hg_set_boundary_flux(self.h, face, a.ctypes.data_as(_dp)),
                    "hg_set_boundary_flux")

    def download(self, field, padded):
        assert padded.flags.c_contiguous and padded.dtype == np.float64
        self._check(self._lib.hg_download(self.h, field, padded.ctypes.data_as(_dp)), "hg_download")

    def timestep(self, time, t_out, t_end, outstyle, endstyle, fixed_dt=-1.0, grav=False):
        dt, ch, cg, dtg = (C.c_double() for _ in range(4))
        self._check(self._lib.hg_timestep(self.h, time, t_out, t_end, outstyle,
                                          endstyle, fixed_dt, C.byref(dt), C.byref(ch),
                                          C.byref(cg), C.byref(dtg)), "hg_timestep")
        if grav:
            return dt.value, ch.value, cg.value, dtg.value
        return dt.value, ch.value

    def gravity(self, time, dt, grvtime, cgrav, cgrav_old, dtgrav):
        """hg_gravity; returns (grvtime, cgrav_old, dtgrav, nsub)."""
        gt, co, dg, ns = C.c_double(grvtime), C.c_double(cgrav_old), C.c_double(dtgrav), \
            C.c_int32(0)
        self._check(self._lib.hg_gravity(self.h, time, dt, C.byref(gt), cgrav, C.byref(co),
                                         C.byref(dg), C.byref(ns)), "hg_gravity")
        return gt.value, co.value, dg.value, ns.value

    def hydro_step(self, dt, ch):
        self._check(self._lib.hg_hydro_step(self.h, dt, ch), "hg_hydro_step")

    def phidamp(self, factor):
        self._check(self._lib.hg_phidamp(self.h, factor), "hg_phidamp")

    def run_loop(self, time, tn, t_out, dt_out, t_end, outstyle, endstyle, tnlim, tn_out,
                 max_steps, damp_factor):
        """Device-resident main loop (hg_run); returns (time, tn, t_out, reason)."""
        t, n, to, r = C.c_double(time), C.c_int64(tn), C.c_double(t_out), C.c_int32(0)
        st = self._lib.hg_run(self.h, C.byref(t), C.byref(n), C.byref(to), dt_out, t_end,
                              outstyle, endstyle, tnlim, tn_out, max_steps, damp_factor,
                              C.byref(r))
        self._check(st, "hg_run", tn=n.value, time=t.value)
        return t.value, n.value, to.value, r.value

    def run_loop_gravity(self, time, tn, t_out, dt_out, t_end, outstyle, endstyle, tnlim,
                         tn_out, max_steps, damp_factor, grvtime, cgrav_old):
        """Device-resident main loop with hyperbolic gravity (hg_run_gravity);
        returns (time, tn, t_out, reason, grvtime, cgrav_old, nsub_total)."""
        t, n, to, r = C.c_double(time), C.c_int64(tn), C.c_double(t_out), C.c_int32(0)
        gt, co, ns = C.c_double(grvtime), C.c_double(cgrav_old), C.c_int64(0)
        st = self._lib.hg_run_gravity(self.h, C.byref(t), C.byref(n), C.byref(to), dt_out,
                                      t_end, outstyle, endstyle, tnlim, tn_out, max_steps,
                                      damp_factor, C.byref(gt), C.byref(co), C.byref(ns),
                                      C.byref(r))
        self._check(st, "hg_run_gravity", tn=n.value, time=t.value)
        return t.value, n.value, to.value, r.value, gt.value, co.value, ns.value

    def set_launch(self, sweep_block_waves: int):
        self._check(self._lib.hg_set_launch(self.h, sweep_block_waves), "hg_set_launch")

    def profile(self, enable=True, kinds=None):
        """Per-launch event timing: every category, or only `kinds`
        (PROF_* values); enable=False turns it off."""
        mode = int(bool(enable))
        if enable and kinds is not None:
            mode = PROF_ONLY | sum(1 << k for k in kinds)
        self._check(self._lib.hg_profile(self.h, mode), "hg_profile")

    def set_fused(self, on=True):
        """Fuse one sweep direction with the RK update (bit-identical):
        0/False off, 1/True the last solved direction (y or z), 2 x."""
        self._check(self._lib.hg_set_fused(self.h, int(on)), "hg_set_fused")

    def time_sweep(self, launches):
        """Milliseconds of `launches` back-to-back sweeps of the current state
        (no per-dispatch events); the state is unchanged."""
        ms = C.c_double()
        self._check(self._lib.hg_time_sweep(self.h, int(launches), C.byref(ms)), "hg_time_sweep")
        return ms.value

    def profile_read(self, kind):
        ms, n, b = C.c_double(), C.c_int64(), C.c_double()
        self._check(self._lib.hg_profile_read(self.h, kind, C.byref(ms), C.byref(n),
                                              C.byref(b)), "hg_profile_read")
        return ms.value, n.value, b.value
