"""Snapshots and restart (the reference's write_bin / readbin / restart,
output.f90:697-795, readbin.f90:16-115, restart.f90:15-64), pinned by a
reference run of Orszag-Tang 64^2 that wrote a snapshot at step 50 and by the
reference's own restart from that snapshot to step 100."""
import hashlib
import os

import numpy as np
import pytest

from hormone_amd import driver as D
from hormone_amd import problems as P
from hormone_amd.snapshot import read_bin, write_bin
from tests.helpers import GOLDEN, NAMES, compare_to_meta, meta

RESTART = dict(n=64, outstyle=2, endstyle=2, tnlim=100, tn_out=50)


def step50():
    z = np.load(os.path.join(GOLDEN, "ot64_step50.npz"))
    snap = {k: z[k].reshape(1, 64, 64) for k in ("d", "v1", "v2", "v3", "e", "b1", "b2", "b3",
                                                 "phi")}
    snap["tn"], snap["time"] = int(z["tn"]), float(z["time"])
    return snap


def test_write_bin_reproduces_the_reference_file():
    """Our writer turns the reference's step-50 state into the very bytes of
    the reference's own snapshot file; the reader reads them back."""
    snap = step50()
    raw = write_bin(snap, mag_on=True)
    m = meta()["ot64_step50"]
    assert len(raw) == m["file_bytes"]
    assert hashlib.sha256(raw).hexdigest() == m["file_sha256"]
    path = "/tmp/hormone_restart_test.bin"
    with open(path, "wb") as f:
        f.write(raw)
    back = read_bin(path, (1, 64, 64), True)
    os.remove(path)
    assert back["tn"] == 50 and back["time"] == snap["time"]
    for k in ("d", "e", "b2", "phi"):
        assert np.array_equal(back[k], snap[k])


def restart_run(engine_cls):
    prm, g, _ = P.make_problem("orszagtang_xy", **RESTART)
    eng = engine_cls(**D.desc_fields(prm, g))
    start = D.load_restart_state(eng, prm, g, step50())
    res = D.run(eng, prm, g, start=start)
    out = D.fetch(eng, g, D.SNAP_FIELDS_MAG)
    return res, {NAMES[i]: out[f] for i, f in enumerate(D.SNAP_FIELDS_MAG)}


def test_oracle_restart_matches_the_reference_restart():
    """Restart from the snapshot (conserve from the primitives read, e_int
    by pressure) and run to step 100: the reference's own restarted run,
    bit for bit (it differs from the uninterrupted run, which the reference
    does not reproduce on restart either)."""
    from oracle.pyoracle import Oracle
    res, state = restart_run(Oracle)
    compare_to_meta(res, state, meta()["ot64_restart50_step100"])
    assert meta()["ot64_restart50_step100"]["time"] != meta()["ot64_step100"]["time"]


@pytest.mark.gpu
def test_gpu_restart_matches_the_reference_restart(hiplib):
    from hormone_amd.native import Hydro
    res, state = restart_run(Hydro)
    compare_to_meta(res, state, meta()["ot64_restart50_step100"])


@pytest.mark.gpu
def test_gpu_uninterrupted_ot64_matches_the_reference_step100(hiplib):
    from hormone_amd.native import Hydro
    from tests.helpers import run_problem
    res, state, _ = run_problem(Hydro, "orszagtang_xy", **RESTART)
    compare_to_meta(res, state, meta()["ot64_step100"])
