"""The C-ABI library loads and exports every entry point the header declares
(CPU-only: no device call is made)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "hormone_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hg_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("hg_create", "hg_destroy", "hg_upload", "hg_download", "hg_timestep",
                 "hg_hydro_step", "hg_phidamp", "hg_last_error", "hg_kat"):
        assert must in names


def test_library_exports_all_declared(hiplib):
    lib = ctypes.CDLL(hiplib._name)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string(hiplib):
    assert b"gfx950" in hiplib.hg_version()


def test_create_rejects_unsupported(hiplib):
    """Descriptor validation happens before any device call."""
    from hormone_amd import native
    from hormone_amd.grid import cartesian_grid
    g = cartesian_grid((8, 1, 1))
    for kw, code in ((dict(bc=(4,) * 12), native.HG_ERR_UNSUPPORTED),
                     (dict(rktype=4), native.HG_ERR_ARG),
                     (dict(bc=(0, 3) + (3,) * 10), native.HG_ERR_ARG),
                     (dict(compswitch=1), native.HG_ERR_UNSUPPORTED),
                     (dict(compswitch=3, spn=2), native.HG_ERR_UNSUPPORTED),
                     (dict(compswitch=2, spn=0), native.HG_ERR_ARG),
                     (dict(compswitch=2, spn=9), native.HG_ERR_ARG)):
        d, keep = native.make_desc(g, **kw)
        h = ctypes.c_void_p()
        assert hiplib.hg_create(ctypes.byref(d), ctypes.byref(h)) == code
        assert not h.value


# The codes each face accepts: the cases of boundarycondition.f90's select
# blocks (x1 inner :54-122 / :124-219, x1 outer :253-347 / :349-493, x2
# :527-598 / :600-692 and :723-787 / :789-877, x3 :911-1004 / :1006-1144 and
# :1175-1265 / :1267-1406); any other code stops the reference.
_ACCEPTED = {
    0: {1, 2, 3, 9, 10}, 1: {1, 2, 3, 4, 5, 9, 10}, 2: {1, 2, 3, 9, 10}, 3: {1, 2, 3, 9, 10},
    4: {1, 2, 3, 4, 5, 9, 10}, 5: {1, 2, 3, 4, 9, 10},                # scalar codes
    6: {1, 2, 3, 9, 10}, 7: {1, 2, 3, 4, 5, 9, 10}, 8: {1, 2, 3, 9, 10}, 9: {1, 2, 3, 9, 10},
    10: {1, 2, 3, 4, 5, 9, 10}, 11: {1, 2, 3, 4, 5, 9, 10},           # vector codes
}


def test_boundary_codes_accepted_per_face(hiplib):
    """hg_create and the oracle accept exactly the codes the reference
    implements on each face (one slot changed at a time from outgoing)."""
    from hormone_amd import native
    from hormone_amd.grid import cartesian_grid
    from oracle.pyoracle import lib as olib
    g = cartesian_grid((8, 8, 8))
    for b in range(12):
        for code in range(1, 12):
            bc = [2] * 12
            bc[b] = code
            for kw in (dict(), dict(compswitch=2, spn=2)):
                d, keep = native.make_desc(g, bc=tuple(bc), **kw)
                h = ctypes.c_void_p()
                st = hiplib.hg_create(ctypes.byref(d), ctypes.byref(h))
                if h.value:
                    hiplib.hg_destroy(h)
                ost = ctypes.c_int(0)
                oh = olib().or_create(ctypes.byref(d), ctypes.byref(ost))
                if oh:
                    olib().or_destroy(oh)
                want = code in _ACCEPTED[b] and not (kw and code >= 9)
                assert (st != native.HG_ERR_UNSUPPORTED) == want, (b, code, kw, st)
                assert (ost.value != native.HG_ERR_UNSUPPORTED) == want, (b, code, kw)
