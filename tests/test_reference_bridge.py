"""The reference-side binding compiles against the unmodified reference.

A scratch copy of the reference's main program gets the maintainer's edits
(hormone_amd/fortran/gpu_bridge.f90 is the module they add; the edits to
hormone.f90 are the five listed in its header), and is compiled with
amdflang against the reference's own module files and objects (built by
oracle/build_ref.py, the same flags), then linked against
libhormone_hip.so.  Compile and link only: the GPU path is not run here.
The bridged program is then run on the CPU path (no &gpucon group) on the
reference's own Sod test problem and must still reproduce the reference's
golden snapshot bit for bit, i.e. the edits leave the reference intact.

Runs only where the reference tree and amdflang exist (the build
container); elsewhere it is skipped.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("HORMONE_REF", "/root/reference")
OBJ = os.path.join(ROOT, "oracle", "_ref", "obj")
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
FLAGS = ["-O3", "-fopenmp", "-ffp-contract=off"]
FORTRAN = os.path.join(ROOT, "hormone_amd", "fortran")
LIB = os.path.join(ROOT, "hormone_amd")

pytestmark = pytest.mark.skipif(
    not (os.path.isdir(os.path.join(REF, "src", "main")) and os.path.exists(FC)
         and os.path.isdir(OBJ) and os.path.exists(os.path.join(LIB, "libhormone_hip.so"))),
    reason="needs the reference tree, amdflang, its build (oracle/build_ref.py) and "
           "libhormone_hip.so")


def bridge_edits(src: str) -> str:
    """The maintainer's five edits to hormone.f90, located by the calls they
    replace (identifiers only, so a re-indented or renumbered file still
    matches): `use gpu_bridge` after the last use statement of the program,
    gpu_init after the first `call pressure`, the four hot-path calls of
    the main loop, gpu_output for output, gpu_sync before the test."""
    lines = src.splitlines()
    out, used, inited, in_loop = [], False, False, False
    last_use = max(i for i, ln in enumerate(lines) if re.match(r"\s*use\s+\w+", ln)
                   and i < next(k for k, x in enumerate(lines) if re.search(r"implicit none", x)))
    for i, ln in enumerate(lines):
        code = ln.split("!")[0]
        if re.match(r"\s*main_loop\s*:\s*do\b", code):
            in_loop = True
        if in_loop and re.match(r"\s*call\s+timestep\s*$", code):
            ln = ln.replace("call timestep", "call gpu_timestep")
        if in_loop:
            ln = re.sub(r"\bcall\s+gravity\b", "call gpu_gravity", ln)
            ln = re.sub(r"\bcall\s+hydro_step\b", "call gpu_hydro_step", ln)
            ln = re.sub(r"\bcall\s+phidamp\b", "call gpu_phidamp", ln)
        ln = re.sub(r"\bcall\s+output\b", "call gpu_output", ln)
        if re.match(r"\s*call\s+test\s*\(", code):
            out.append("    call gpu_sync")
        out.append(ln)
        if i == last_use:
            out.append("  use gpu_bridge")
            used = True
        if not inited and re.match(r"\s*call\s+pressure\s*$", code):
            out.append("  call gpu_init")
            inited = True
    assert used and inited and in_loop
    text = "\n".join(out) + "\n"
    for name in ("gpu_timestep", "gpu_gravity", "gpu_hydro_step", "gpu_phidamp", "gpu_output",
                 "gpu_sync"):
        assert f"call {name}" in text, name
    return text


@pytest.fixture(scope="module")
def bridged(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("bridge")
    mod = tmp / "mod"
    mod.mkdir()
    with open(os.path.join(REF, "src", "main", "hormone.f90")) as f:
        (tmp / "hormone.f90").write_text(bridge_edits(f.read()))
    objs = []
    for src in ("hormone_hip_mod.f90", "gpu_bridge.f90"):
        o = tmp / (src[:-4] + ".o")
        subprocess.run([FC, *FLAGS, "-I", OBJ, "-J", str(mod), "-c",
                        os.path.join(FORTRAN, src), "-o", str(o)], check=True)
        objs.append(str(o))
    o = tmp / "hormone.o"
    subprocess.run([FC, *FLAGS, "-I", OBJ, "-I", str(mod), "-c", str(tmp / "hormone.f90"),
                    "-o", str(o)], check=True)
    ref_objs = sorted(os.path.join(OBJ, f) for f in os.listdir(OBJ)
                      if f.endswith(".o") and f != "hormone.o")
    exe = tmp / "hormone_gpu"
    subprocess.run([FC, *FLAGS, str(o), *objs, *ref_objs, "-L", LIB, "-lhormone_hip",
                    f"-Wl,-rpath,{LIB}", "-o", str(exe)], check=True)
    return tmp, exe


def test_bridge_compiles_and_links_against_the_reference(bridged):
    tmp, exe = bridged
    assert exe.exists()
    syms = subprocess.run(["nm", "-D", "--undefined-only", str(exe)], capture_output=True,
                          text=True, check=True).stdout
    for s in ("hg_create", "hg_timestep", "hg_gravity", "hg_hydro_step", "hg_phidamp",
              "hg_upload", "hg_download"):
        assert re.search(rf"\b{s}\b", syms), f"{s} is not bound"


def test_bridged_reference_without_gpucon_is_the_reference(bridged):
    """No &gpucon group: every gpu_* call is the reference's routine, and the
    reference's own sodshock_x test run reproduces its golden snapshot."""
    tmp, exe = bridged
    run = tmp / "run"
    (run / "work" / "data").mkdir(parents=True)
    for d in ("para", "tests"):
        os.symlink(os.path.join(REF, d), run / d)
    (run / "work" / "startfile").write_text(
        "&startcon\n start=0\n simtype='sodshock_x test'\n parafile='parameters'\n/\n")
    (run / "work" / "parameters").write_text("")
    r = subprocess.run([str(exe)], cwd=run / "work", capture_output=True, text=True,
                       timeout=600, env={**os.environ, "OMP_NUM_THREADS": "2"})
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "sodshock_x test: passed" in r.stdout, r.stdout[-2000:]
    assert "Identical!" in r.stdout, r.stdout[-2000:]
