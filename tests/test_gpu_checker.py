"""GPU answers written with sk_write_solution and judged by the independent answer checker
(oracle/checker.py, held to the reference checker's verdicts in tests/test_checker.py)."""
from pathlib import Path

import pytest

from oracle import checker
from tests.mps_ref import read_mps

pytestmark = pytest.mark.gpu
NETLIB = Path(__file__).resolve().parent / "golden" / "netlib"


def to_lp(sk, d):
    return sk.Lp.from_arrays(d["num_rows"], d["num_cols"], d["col_starts"], d["row_indices"],
                             d["values"], d["cost"], d["col_lower"], d["col_upper"],
                             d["row_lower"], d["row_upper"], d.get("sense", 1), d.get("offset", 0.0))


@pytest.mark.parametrize("name", ["afiro", "sc50a", "sc50b", "blend"])
def test_fp64_engine_answers_pass_the_checker(gpu, tmp_path, name):
    # the reference stop rule at a tight tolerance: the engine claims `optimal`, so the
    # checker holds the point to every LP optimality test (feasibility 1e-7, duality 1e-9)
    d = read_mps(NETLIB / f"{name}.mps")
    lp = to_lp(gpu, d)
    r = gpu.solve(lp, pdhg_tolerance=1e-9, pdhg_precision="fp64", iteration_limit=400000)
    assert r.status_name == "optimal", r.info["message"]
    sol = tmp_path / f"{name}.sol"
    gpu.write_solution(sol, lp, r, model_name=name, col_names=d["col_names"],
                           row_names=d["row_names"])
    report = checker.verify(d, checker.read_sol(sol))
    assert report.verified, [line for line in report.lines if not line[0]]


def test_fp32_answer_is_checked_on_what_it_claims(gpu, tmp_path):
    # a relative-KKT 1e-4 answer from FP32 iterates claims `feasible`, not `optimal`: the
    # checker recomputes its objective and activities (which must agree with the file) and
    # skips the optimality tests, exactly as for the reference's own `feasible` answers
    lp = gpu.Lp.generate(3, 5)
    d = lp.export()
    r = gpu.solve(lp, pdhg_tolerance=1e-4, pdhg_stop_at_request=True, pdhg_absolute_primal=False)
    assert r.converged
    sol = tmp_path / "c3_5.sol"
    gpu.write_solution(sol, lp, r, model_name="c3_5")
    report = checker.verify(d, checker.read_sol(sol), primal_tol=1e-3)
    outcome = dict(report.outcome())
    assert outcome["structure"] and outcome["objective"] and outcome["activity agreement"]
    assert "strong duality" not in outcome  # not an optimality claim
