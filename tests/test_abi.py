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
