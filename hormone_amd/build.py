"""Build the HIP library ``hormone_amd/libhormone_hip.so`` for gfx950.

FP64 parity with the reference requires no FMA contraction
(-ffp-contract=off) and IEEE division/square root (no fast-math).  Packed
FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libhormone_hip.so")
SOURCES = ["hg_kernels.hip", "hg_api.hip"]
ARCH = os.environ.get("HORMONE_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags():
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-ffp-contract=off", "-fno-fast-math",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function"]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *flags(), "-shared", "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, OUT)
    return OUT


FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "hg_driver")


def build_fortran(verbose: bool = False) -> str | None:
    """The iso_c_binding interface module and the Fortran host program,
    linked against libhormone_hip.so (rpath $ORIGIN/..).  Needs amdflang."""
    fc = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
    if not os.path.exists(fc):
        return None
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        obj = os.path.join(tmp, "hormone_hip_mod.o")
        cmds = [
            [fc, "-O2", "-ffp-contract=off", "-c", os.path.join(FORTRAN, "hormone_hip_mod.f90"),
             "-o", obj, "-J", tmp],
            [fc, "-O2", "-ffp-contract=off", "-I", tmp, os.path.join(FORTRAN, "hg_driver.f90"),
             obj, "-L", HERE, "-lhormone_hip", "-Wl,-rpath,$ORIGIN/..", "-o", DRIVER + ".tmp"],
        ]
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                sys.stderr.write(r.stdout + r.stderr)
                raise RuntimeError("amdflang failed")
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
