"""The CPU oracle against the reference's own golden snapshots (bit-exact).

Pins oracle/hormone_oracle.c against tests/NAME.bin of the reference
(converted to tests/golden/*.npz and golden_meta.json by
tests/golden/make_fixtures.py) and against reference runs at the benchmark
sizes.  Runs on the CPU.
"""
import pytest

from oracle.pyoracle import Oracle
from tests.helpers import compare_to_meta, compare_to_npz, meta, run_problem

ONE_D = [f"{k}_{d}" for k in ("sodshock", "briowushock") for d in "xyz"]


@pytest.mark.parametrize("flat", [False, True], ids=["muscl", "flat"])
@pytest.mark.parametrize("name", ONE_D)
def test_oracle_1d_goldens(name, flat):
    res, state, _ = run_problem(Oracle, name, flat=flat)
    compare_to_npz(res, state, name + ("_flat" if flat else ""))


@pytest.mark.parametrize("flat", [False, True], ids=["muscl", "flat"])
@pytest.mark.parametrize("plane", ["xy", "xz", "yz"])
def test_oracle_orszagtang_goldens(plane, flat):
    name = f"orszagtang_{plane}"
    res, state, _ = run_problem(Oracle, name, flat=flat)
    compare_to_meta(res, state, meta()[name + ("_flat" if flat else "")])


def test_oracle_sod1024_courant04():
    """Benchmark-size Sod (1024 cells, courant 0.4) vs a reference run."""
    res, state, _ = run_problem(Oracle, "sodshock_x", n=1024, courant=0.4)
    compare_to_npz(res, state, "sod1024_c04")


def test_oracle_ot512_prefix():
    """Orszag-Tang 512^2, first 20 steps, vs a reference run."""
    res, state, _ = run_problem(Oracle, "orszagtang_xy", n=512, outstyle=2, endstyle=2,
                                tnlim=20)
    compare_to_meta(res, state, meta()["ot512_20steps"])
