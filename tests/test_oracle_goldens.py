"""The CPU oracle against the reference's own golden snapshots (bit-exact).

Pins oracle/hormone_oracle.c against tests/NAME.bin of the reference
(converted to tests/golden/*.npz and golden_meta.json by
tests/golden/make_fixtures.py) and against reference runs at the benchmark
sizes.  Runs on the CPU.
"""
import pytest

from oracle.pyoracle import Oracle
from tests.helpers import (EXTRA_CASES, compare_to_meta, compare_to_npz, meta, run_extra,
                           run_problem)

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


@pytest.mark.parametrize("case", list(EXTRA_CASES))
def test_oracle_extra_reference_runs(case):
    """Reflective and outgoing boundaries, minmod, RK2 and RK1 against runs of
    the reference program built here (tests/golden/make_fixtures.py)."""
    res, state, _ = run_extra(Oracle, case)
    compare_to_meta(res, state, meta()[case])


def test_oracle_reflective_briowu_stops_like_the_reference():
    """Reflective walls flip the uniform normal field of Brio-Wu; the
    reference stops during step 6 with a negative internal energy at the
    wall cell.  The oracle stops at the same step and cell."""
    from oracle.pyoracle import HgError
    fail = meta()["briowu_x_reflective_failure"]
    name, kw = EXTRA_CASES["briowu_x_reflective_5steps"]
    kw = dict(kw, outstyle=1, endstyle=1, t_end=300.0, dt_out=300.0)
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    prm, g, st = P.make_problem(name, **kw)
    eng = Oracle(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    res = D.run(eng, prm, g, max_steps=fail["completed_steps"])
    assert res.tn == fail["completed_steps"]
    with pytest.raises(HgError) as ei:
        D.run(eng, prm, g, max_steps=1)
    assert list(ei.value.cell) == fail["cell"]


def test_oracle_linear_boundary_stops_like_the_reference():
    """Orszag-Tang yz with linear codes at the x3 boundaries: the reference
    stops during step 50 with a negative internal energy next to the x3
    inner boundary; the oracle stops at the same step and cell."""
    from hormone_amd import driver as D
    from hormone_amd import problems as P
    from oracle.pyoracle import HgError
    fail = meta()["ot_yz_linear_failure"]
    name, kw = EXTRA_CASES["ot_yz_linear"]
    prm, g, st = P.make_problem(name, **dict(kw, tnlim=100, tn_out=100))
    eng = Oracle(**D.desc_fields(prm, g))
    D.load_initial_state(eng, prm, g, st)
    res = D.run(eng, prm, g, max_steps=fail["completed_steps"])
    assert res.tn == fail["completed_steps"]
    with pytest.raises(HgError) as ei:
        D.run(eng, prm, g, max_steps=1)
    assert list(ei.value.cell) == fail["cell"]


@pytest.mark.parametrize("flat", [False, True], ids=["muscl", "flat"])
def test_oracle_kh_goldens(flat):
    """2D hydro Kelvin-Helmholtz with two advected species (compswitch 2):
    the reference's KHinstability goldens, every field and both species."""
    res, state, _ = run_problem(Oracle, "KHinstability", flat=flat)
    compare_to_meta(res, state, meta()["KHinstability" + ("_flat" if flat else "")])
