#!/usr/bin/env python3
"""Build recipe for the reference HORMONE Fortran program, used only as a checker.

Compiles the reference's own Fortran sources where they lie (read-only) with
ROCm's ``amdflang`` into ``oracle/_ref/`` (git-ignored): objects and module
files under ``oracle/_ref/obj`` and the executable ``oracle/_ref/hormone``.
Serial + OpenMP build (no MPI, no PETSc), ``-O3`` without FMA contraction,
which reproduces the reference's committed gfortran golden snapshots bit for
bit.  The compile order is derived here from the ``module``/``use``
statements; the reference's own Makefile is not used.

Nothing in the product imports this file or what it builds; the fixture
generator under ``tests/golden/`` runs the resulting binary in this container.
"""
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

REF = os.environ.get("HORMONE_REF", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
OBJ = os.path.join(OUT, "obj")
FC = os.environ.get("FC", "amdflang")
FLAGS = ["-O3", "-fopenmp", "-ffp-contract=off"]

MOD_RE = re.compile(r"^\s*module\s+(\w+)\s*(!.*)?$", re.I)
USE_RE = re.compile(r"^\s*use\s*(?:,\s*intrinsic\s*::)?\s*(\w+)", re.I)


def sources():
    main = os.path.join(REF, "src", "main")
    init = os.path.join(REF, "src", "init")
    files = []
    for d in (main, init):
        for f in sorted(os.listdir(d)):
            if f.endswith((".f90", ".F90")):
                files.append(os.path.join(d, f))
    return files


def scan(path):
    provides, uses = set(), set()
    with open(path, errors="replace") as fh:
        for line in fh:
            m = MOD_RE.match(line)
            if m and m.group(1).lower() != "procedure":
                provides.add(m.group(1).lower())
            u = USE_RE.match(line)
            if u:
                uses.add(u.group(1).lower())
    return provides, uses


def levels(files):
    info = {f: scan(f) for f in files}
    owner = {}
    for f, (prov, _) in info.items():
        for m in prov:
            owner[m] = f
    deps = {f: {owner[u] for u in uses if u in owner and owner[u] != f}
            for f, (_, uses) in info.items()}
    done, out = set(), []
    while len(done) < len(files):
        lvl = [f for f in files if f not in done and deps[f] <= done]
        if not lvl:
            raise RuntimeError("cyclic module dependencies: %s" %
                               sorted(set(files) - done))
        out.append(lvl)
        done.update(lvl)
    return out


def compile_one(src):
    obj = os.path.join(OBJ, os.path.splitext(os.path.basename(src))[0] + ".o")
    cmd = [FC, *FLAGS, "-c", src, "-o", obj, "-J", OBJ]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + src)
    return obj


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not found at " + REF)
    os.makedirs(OBJ, exist_ok=True)
    files = sources()
    objs = []
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for lvl in levels(files):
            objs += list(ex.map(compile_one, lvl))
    exe = os.path.join(OUT, "hormone")
    subprocess.run([FC, *FLAGS, "-o", exe, *objs], check=True)
    print(exe)


if __name__ == "__main__":
    main()
