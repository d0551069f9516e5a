#!/usr/bin/env python3
"""Generate the golden fixtures under tests/golden/ (run in the build
container, where the reference tree is mounted read-only).

Two sources, both data produced by the reference itself:

1. The reference's own committed golden snapshots (tests/NAME.bin) for the
   Cartesian hydro/MHD problems, read with hormone_amd.snapshot.read_bin and
   re-stored in our format: small 1D states as float64 arrays, 2D states as a
   SHA-256 digest of each field's little-endian bytes plus norms.
2. Runs of the reference program built from source by oracle/build_ref.py
   (oracle/_ref/hormone) at the benchmark sizes: Sod 1024 cells at
   courant 0.4 to t=0.2, and Orszag-Tang 512^2 (a 20-step prefix and the
   full run to t=0.5).  The binary runs in a scratch directory under /tmp
   whose ../para points at the reference's parameter files.

Usage: python tests/golden/make_fixtures.py [--skip-long | --extra-only | --kh-only | --restart-only]
"""
from __future__ import annotations

import glob
import hashlib
import json
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from hormone_amd.snapshot import read_bin  # noqa: E402

REF = os.environ.get("HORMONE_REF", "/root/reference")
BIN = os.path.join(ROOT, "oracle", "_ref", "hormone")
FIELDS = ("d", "v1", "v2", "v3", "e")
FIELDS_MAG = FIELDS + ("b1", "b2", "b3", "phi")


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def summary(snap: dict, names) -> dict:
    out = {"tn": snap["tn"], "time": snap["time"], "fields": {}}
    for nm in names:
        a = snap[nm]
        out["fields"][nm] = {"sha256": digest(a), "sum": float(a.sum()),
                             "abs_sum": float(np.abs(a).sum())}
    return out


def shape_of(n):
    return (n[2], n[1], n[0])


CASES_1D = {
    "sodshock": (False, 500), "briowushock": (True, 500)}


def convert_reference_goldens(meta: dict):
    for kind, (mag, n) in CASES_1D.items():
        for d in "xyz":
            for flat in ("", "_flat"):
                name = f"{kind}_{d}{flat}"
                nn = tuple(n if c == d else 1 for c in "xyz")
                snap = read_bin(os.path.join(REF, "tests", name + ".bin"), shape_of(nn), mag)
                names = FIELDS_MAG if mag else FIELDS
                np.savez_compressed(os.path.join(HERE, name + ".npz"), tn=snap["tn"],
                                    time=snap["time"],
                                    **{k: snap[k].reshape(-1) for k in names})
                meta[name] = summary(snap, names)
                meta[name]["source"] = f"reference tests/{name}.bin"
    for plane in ("xy", "xz", "yz"):
        for flat in ("", "_flat"):
            name = f"orszagtang_{plane}{flat}"
            nn = tuple(128 if c in plane else 1 for c in "xyz")
            snap = read_bin(os.path.join(REF, "tests", name + ".bin"), shape_of(nn), True)
            meta[name] = summary(snap, FIELDS_MAG)
            meta[name]["source"] = f"reference tests/{name}.bin"
    convert_kh_goldens(meta)


def convert_kh_goldens(meta: dict):
    """KHinstability{,_flat}: 2D hydro, 128^2, two advected species."""
    for flat in ("", "_flat"):
        name = f"KHinstability{flat}"
        snap = read_bin(os.path.join(REF, "tests", name + ".bin"), (1, 128, 128), False,
                        compswitch=2)
        meta[name] = summary(snap, FIELDS + ("spc1", "spc2"))
        meta[name]["species"] = snap["species"]
        meta[name]["source"] = f"reference tests/{name}.bin"


def run_reference(simtype: str, params: str, threads: int = 8) -> str:
    run = "/tmp/hormone_fixture_run"
    shutil.rmtree(run, ignore_errors=True)
    os.makedirs(os.path.join(run, "work", "data"))
    os.symlink(os.path.join(REF, "para"), os.path.join(run, "para"))
    os.symlink(os.path.join(REF, "tests"), os.path.join(run, "tests"))
    with open(os.path.join(run, "work", "startfile"), "w") as f:
        f.write(f"&startcon\n start=0\n simtype='{simtype}'\n parafile='parameters'\n/\n")
    with open(os.path.join(run, "work", "parameters"), "w") as f:
        f.write(params)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    with open(os.path.join(run, "log.txt"), "w") as log:
        subprocess.run([BIN], cwd=os.path.join(run, "work"), env=env, check=True,
                       stdout=log, stderr=subprocess.STDOUT)
    bins = glob.glob(os.path.join(run, "work", "data", "bin*.dat"))
    # the final snapshot is the one with the largest step count
    return max(bins, key=lambda b: read_header(b)[0])


def restart_reference(simtype: str, params: str, start: int, threads: int = 8) -> str:
    """Restart the reference program in the scratch directory of the last
    run_reference from its snapshot at step `start` (startfile start=N,
    restart.f90:15-64); returns the last snapshot written."""
    run = "/tmp/hormone_fixture_run"
    with open(os.path.join(run, "work", "startfile"), "w") as f:
        f.write(f"&startcon\n start={start}\n simtype='{simtype}'\n parafile='parameters'\n/\n")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    before = set(glob.glob(os.path.join(run, "work", "data", "bin*.dat")))
    with open(os.path.join(run, "log_restart.txt"), "w") as log:
        subprocess.run([BIN], cwd=os.path.join(run, "work"), env=env, check=True,
                       stdout=log, stderr=subprocess.STDOUT)
    bins = glob.glob(os.path.join(run, "work", "data", "bin*.dat"))
    return max(bins, key=lambda b: read_header(b)[0])


RESTART_PARAMS = ("&gridcon ie=64 je=64 /\n&out_con outstyle=2 endstyle=2 tnlim=100 "
                  "tn_out=50 output_ascii=.false. /\n")


def reference_restart_runs(meta: dict):
    """Orszag-Tang 64^2: a run to step 100 with snapshots every 50 steps,
    then a restart from the step-50 snapshot to step 100.  Stored: the
    step-50 snapshot (input of the restart) as arrays plus its file digest,
    and digests of both step-100 states (uninterrupted and restarted)."""
    b100 = run_reference("orszagtang_xy", RESTART_PARAMS)
    snap100 = read_bin(b100, (1, 64, 64), True)
    b50 = [b for b in glob.glob(os.path.join(os.path.dirname(b100), "bin*.dat"))
           if read_header(b)[0] == 50][0]
    raw50 = open(b50, "rb").read()
    snap50 = read_bin(b50, (1, 64, 64), True)
    np.savez_compressed(os.path.join(HERE, "ot64_step50.npz"), tn=snap50["tn"],
                        time=snap50["time"], **{k: snap50[k].reshape(-1) for k in FIELDS_MAG})
    meta["ot64_step50"] = summary(snap50, FIELDS_MAG)
    meta["ot64_step50"]["file_sha256"] = hashlib.sha256(raw50).hexdigest()
    meta["ot64_step50"]["file_bytes"] = len(raw50)
    meta["ot64_step50"]["source"] = "reference run: orszagtang_xy ie=je=64, snapshot at tn=50"
    meta["ot64_step100"] = summary(snap100, FIELDS_MAG)
    meta["ot64_step100"]["source"] = "reference run: orszagtang_xy ie=je=64, tnlim=100"
    br = restart_reference("orszagtang_xy", RESTART_PARAMS, 50)
    snapr = read_bin(br, (1, 64, 64), True)
    meta["ot64_restart50_step100"] = summary(snapr, FIELDS_MAG)
    meta["ot64_restart50_step100"]["source"] = (
        "reference run: orszagtang_xy ie=je=64 restarted (start=50) from its tn=50 snapshot, to tnlim=100")


def read_header(path):
    raw = open(path, "rb").read(20)
    return int(np.frombuffer(raw, "<i4", 1, 4)[0]), float(np.frombuffer(raw, "<f8", 1, 8)[0])


def reference_runs(meta: dict, long: bool):
    meta.pop("ot512_t05", None)
    # Sod 1024 cells, courant 0.4, to t = 0.2 (200 ms)
    p = ("&gridcon ie=1024 /\n&out_con output_ascii=.false. /\n"
         "&simucon courant=0.4d0 /\n")
    b = run_reference("sodshock_x", p)
    snap = read_bin(b, (1, 1, 1024), False)
    np.savez_compressed(os.path.join(HERE, "sod1024_c04.npz"), tn=snap["tn"], time=snap["time"],
                        **{k: snap[k].reshape(-1) for k in FIELDS})
    meta["sod1024_c04"] = summary(snap, FIELDS)
    meta["sod1024_c04"]["source"] = "reference run: sodshock_x ie=1024 courant=0.4"
    # Orszag-Tang 512^2, first 20 steps
    p = ("&gridcon ie=512 je=512 /\n&out_con outstyle=2 endstyle=2 tnlim=20 tn_out=20 "
         "output_ascii=.false. /\n")
    b = run_reference("orszagtang_xy", p)
    snap = read_bin(b, (1, 512, 512), True)
    meta["ot512_20steps"] = summary(snap, FIELDS_MAG)
    meta["ot512_20steps"]["source"] = "reference run: orszagtang_xy ie=je=512, tnlim=20"
    if long:
        # The full run to t=0.5 stops inside the reference: internalenergy
        # finds e_int <= 0 (eos.f90:833-839).  Record where, and the state
        # 50 steps before.
        p = ("&gridcon ie=512 je=512 /\n&out_con output_ascii=.false. /\n")
        try:
            run_reference("orszagtang_xy", p)
            fail = None
        except subprocess.CalledProcessError:
            fail = True
        log = open("/tmp/hormone_fixture_run/log.txt").read()
        import re
        steps = [int(m) for m in re.findall(r"tn =\s+(\d+)", log)]
        cell = re.search(r"i=\s*(\d+), j=\s*(\d+), k=\s*(\d+)", log)
        meta["ot512_t05_failure"] = {
            "stops": bool(fail or "Negative internal energy" in log),
            "message": "Error in internalenergy, Negative internal energy",
            "completed_steps": steps[-1] if steps else None,
            "cell": [int(cell.group(q)) for q in (1, 2, 3)] if cell else None,
            "source": "reference run: orszagtang_xy ie=je=512 to t=0.5 (stops early)"}
        p = ("&gridcon ie=512 je=512 /\n&out_con outstyle=2 endstyle=2 tnlim=2200 "
             "tn_out=2200 output_ascii=.false. /\n")
        b = run_reference("orszagtang_xy", p)
        snap = read_bin(b, (1, 512, 512), True)
        meta["ot512_2200steps"] = summary(snap, FIELDS_MAG)
        meta["ot512_2200steps"]["source"] = "reference run: orszagtang_xy ie=je=512, tnlim=2200"


# Reference runs that exercise what the committed goldens do not: boundary
# codes 1 (reflective) and 2 (outgoing) on each axis, the minmod limiter,
# and RK orders 2 and 1.  Each entry: simtype, the Python problem keywords
# (tests/helpers.EXTRA_CASES restates them), and the namelist overrides the
# reference reads on top of para/parameters_base and para/parameters_SIMTYPE.
def _bc_text(axis, code):
    return (f"&bouncon bc{axis}is={code} bc{axis}os={code} bc{axis}iv={code} "
            f"bc{axis}ov={code} /\n")


EXTRA = {
    "sod_x_reflective": ("sodshock_x", "&gridcon ie=256 /\n" + _bc_text(1, 1)
                         + "&out_con t_end=400d0 dt_out=400d0 output_ascii=.false. /\n"),
    "sod_x_outgoing": ("sodshock_x", "&gridcon ie=256 /\n" + _bc_text(1, 2)
                       + "&out_con t_end=300d0 dt_out=300d0 output_ascii=.false. /\n"),
    "sod_y_outgoing": ("sodshock_y", "&gridcon je=256 /\n" + _bc_text(2, 2)
                       + "&out_con t_end=300d0 dt_out=300d0 output_ascii=.false. /\n"),
    # reflective walls reverse the normal B of the uniform B_x = 0.75: the
    # reference stops during step 6 (negative e_int at the wall); the
    # 5-step state and the stop are both pinned
    "briowu_x_reflective_5steps": ("briowushock_x", "&gridcon ie=256 /\n" + _bc_text(1, 1)
                                   + "&out_con outstyle=2 endstyle=2 tnlim=5 tn_out=5 "
                                     "output_ascii=.false. /\n"),
    "briowu_x_reflective_failure": ("briowushock_x", "&gridcon ie=256 /\n" + _bc_text(1, 1)
                                    + "&out_con t_end=300d0 dt_out=300d0 "
                                      "output_ascii=.false. /\n"),
    "briowu_z_outgoing": ("briowushock_z", "&gridcon ke=256 /\n" + _bc_text(3, 2)
                          + "&out_con t_end=300d0 dt_out=300d0 output_ascii=.false. /\n"),
    "sod_x_minmod_rk2": ("sodshock_x", "&gridcon ie=256 /\n&simucon flux_limiter='minmod' "
                         "rktype=2 /\n&out_con output_ascii=.false. /\n"),
    "briowu_y_minmod": ("briowushock_y", "&gridcon je=256 /\n&simucon flux_limiter='minmod' /\n"
                        "&out_con output_ascii=.false. /\n"),
    "sod_x_flat_rk1": ("sodshock_x", "&gridcon ie=256 /\n&simucon flux_limiter='flat' "
                       "rktype=1 /\n&out_con output_ascii=.false. /\n"),
    "ot64_minmod_rk2": ("orszagtang_xy", "&gridcon ie=64 je=64 /\n&simucon flux_limiter='minmod' "
                        "rktype=2 /\n&out_con outstyle=2 endstyle=2 tnlim=100 tn_out=100 "
                        "output_ascii=.false. /\n"),
}


# Linear boundary codes 4/5 where the reference implements them (x1 outer,
# x3 inner, x3 outer).  Orszag-Tang in the yz plane has nonzero v2 and b2 at
# the x3 inner boundary, where the reference leaves v2 and b2 of the outer
# ghost unset (boundarycondition.f90:1081,1088,1104,1114).
_OUT = "&out_con t_end=300d0 dt_out=300d0 output_ascii=.false. /\n"
# Brio-Wu along z with linear + outgoing codes at x3 inner turns to NaN in
# the reference after about 600 steps (the x3 outer inflow grows without
# bound); the pinned state is at 450, after both boundaries have acted
LINEAR_BRIOWU_Z_STEPS = 450
LINEAR = {
    "sod_x_linear": ("sodshock_x", "&gridcon ie=256 /\n&bouncon bc1is=2 bc1iv=2 bc1os=4 "
                     "bc1ov=4 /\n" + _OUT),
    "briowu_x_linear5": ("briowushock_x", "&gridcon ie=256 /\n&bouncon bc1is=3 bc1iv=3 "
                         "bc1os=5 bc1ov=5 /\n" + _OUT),
    "briowu_z_linear5": ("briowushock_z", "&gridcon ke=256 /\n&bouncon bc3is=5 bc3iv=5 "
                         "bc3os=4 bc3ov=4 /\n&out_con outstyle=2 endstyle=2 tnlim=TN "
                         "tn_out=TN output_ascii=.false. /\n"),
    # the run stops during step 50 (negative e_int next to the x3 inner
    # boundary): the state after 40 steps and the stop are both pinned
    "ot_yz_linear": ("orszagtang_yz", "&gridcon je=64 ke=64 /\n&bouncon bc3is=4 bc3iv=4 "
                     "bc3os=4 bc3ov=5 /\n&out_con outstyle=2 endstyle=2 tnlim=40 "
                     "tn_out=40 output_ascii=.false. /\n"),
    "ot_yz_linear_failure": ("orszagtang_yz", "&gridcon je=64 ke=64 /\n&bouncon bc3is=4 "
                             "bc3iv=4 bc3os=4 bc3ov=5 /\n&out_con outstyle=2 endstyle=2 "
                             "tnlim=100 tn_out=100 output_ascii=.false. /\n"),
}


def reference_runs_linear(meta: dict):
    import re
    for name, (simtype, params) in LINEAR.items():
        params = params.replace("TN", str(LINEAR_BRIOWU_Z_STEPS))
        b = run_reference(simtype, params)
        if name.endswith("_failure"):
            log = open("/tmp/hormone_fixture_run/log.txt").read()
            steps = [int(m) for m in re.findall(r"tn =\s+(\d+)", log)]
            cell = re.search(r"i=\s*(\d+), j=\s*(\d+), k=\s*(\d+)", log)
            meta[name] = {
                "stops": "Negative internal energy" in log,
                "completed_steps": steps[-1] if steps else None,
                "cell": [int(cell.group(q)) for q in (1, 2, 3)] if cell else None,
                "source": f"reference run: {simtype} with " + " ".join(params.split())}
            continue
        if simtype.startswith("orszagtang"):
            shape = (64, 64, 1)
        else:
            shape = {"x": (1, 1, 256), "y": (1, 256, 1), "z": (256, 1, 1)}[simtype[-1]]
        mag = not simtype.startswith("sodshock")
        snap = read_bin(b, shape, mag)
        meta[name] = summary(snap, FIELDS_MAG if mag else FIELDS)
        meta[name]["source"] = f"reference run: {simtype} with " + " ".join(params.split())
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), tn=snap["tn"], time=snap["time"],
                            **{k: snap[k].reshape(-1) for k in (FIELDS_MAG if mag else FIELDS)})


def reference_grid_file(meta: dict):
    """gridfile.bin the reference writes at start-up (write_grid_bin,
    output.f90:432-452) for Orszag-Tang 64^2."""
    run_reference("orszagtang_xy", "&gridcon ie=64 je=64 /\n&out_con outstyle=2 endstyle=2 "
                  "tnlim=1 tn_out=1 output_ascii=.false. /\n")
    # dx(is-2) of each axis is never set by gridset (gridset.f90:54-59 starts
    # at is-1), so the file holds whatever that memory held: keep the file
    raw = open("/tmp/hormone_fixture_run/work/data/gridfile.bin", "rb").read()
    with open(os.path.join(HERE, "gridfile_ot64.bin"), "wb") as f:
        f.write(raw)
    meta["gridfile_ot64"] = {"file_bytes": len(raw),
                             "source": "reference run: orszagtang_xy with ie=64 je=64, "
                                       "data/gridfile.bin (tests/golden/gridfile_ot64.bin)"}


def reference_runs_extra(meta: dict):
    import re
    meta.pop("briowu_x_reflective", None)
    for name, (simtype, params) in EXTRA.items():
        b = run_reference(simtype, params)
        if name.endswith("_failure"):
            log = open("/tmp/hormone_fixture_run/log.txt").read()
            steps = [int(m) for m in re.findall(r"tn =\s+(\d+)", log)]
            cell = re.search(r"i=\s*(\d+), j=\s*(\d+), k=\s*(\d+)", log)
            meta[name] = {
                "stops": "Negative internal energy" in log,
                "message": "Error in internalenergy, Negative internal energy",
                "completed_steps": steps[-1] if steps else None,
                "cell": [int(cell.group(q)) for q in (1, 2, 3)] if cell else None,
                "source": f"reference run: {simtype} with " + " ".join(params.split())}
            continue
        mag = simtype.startswith(("briowushock", "orszagtang"))
        if simtype.startswith("orszagtang"):
            shape = (1, 64, 64)
        else:
            ax = simtype[-1]
            shape = {"x": (1, 1, 256), "y": (1, 256, 1), "z": (256, 1, 1)}[ax]
        snap = read_bin(b, shape, mag)
        meta[name] = summary(snap, FIELDS_MAG if mag else FIELDS)
        meta[name]["source"] = f"reference run: {simtype} with " + " ".join(params.split())


def main():
    long = "--skip-long" not in sys.argv
    if "--restart-only" in sys.argv:
        path = os.path.join(HERE, "golden_meta.json")
        meta = json.load(open(path))
        reference_restart_runs(meta)
        with open(path, "w") as f:
            json.dump(meta, f, indent=1, sort_keys=True)
        print("wrote", path)
        return
    if "--kh-only" in sys.argv:
        path = os.path.join(HERE, "golden_meta.json")
        meta = json.load(open(path))
        convert_kh_goldens(meta)
        with open(path, "w") as f:
            json.dump(meta, f, indent=1, sort_keys=True)
        print("wrote", path)
        return
    if "--linear-only" in sys.argv:
        path = os.path.join(HERE, "golden_meta.json")
        meta = json.load(open(path))
        reference_runs_linear(meta)
        reference_grid_file(meta)
        with open(path, "w") as f:
            json.dump(meta, f, indent=1, sort_keys=True)
        print("wrote", path)
        return
    if "--extra-only" in sys.argv:
        path = os.path.join(HERE, "golden_meta.json")
        meta = json.load(open(path))
        reference_runs_extra(meta)
        with open(path, "w") as f:
            json.dump(meta, f, indent=1, sort_keys=True)
        print("wrote", path)
        return
    path = os.path.join(HERE, "golden_meta.json")
    meta = json.load(open(path)) if os.path.exists(path) else {}
    convert_reference_goldens(meta)
    reference_runs(meta, long)
    reference_runs_extra(meta)
    reference_runs_linear(meta)
    reference_grid_file(meta)
    reference_restart_runs(meta)
    with open(path, "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()
