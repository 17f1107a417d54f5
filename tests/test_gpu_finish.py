"""Warm starts and the FP64 finish of LPs the interior point declines (sk_batch_warm_start,
sk_pdhg_refine): a NaN warm start is the cold start bit for bit; a start at the answer stops at
the first evaluation; and the 2,000,000 x 5,000,000 power-law LP made optimal at a drawn point
(tests/lp_cases.kkt_backwards_from) goes FP32 PDHG at 1e-4 -> FP64 finish -> its known optimum
within 1e-6 relative objective, as the oracle measures it and the answer checker verifies."""
import numpy as np
import pytest

from oracle import checker
from tests.lp_cases import kkt_backwards_from

pytestmark = pytest.mark.gpu

REL = dict(pdhg_stop_at_request=True, pdhg_absolute_primal=False)


def to_lp(sk, d):
    return sk.Lp.from_arrays(d["num_rows"], d["num_cols"], d["col_starts"], d["row_indices"],
                             d["values"], d["cost"], d["col_lower"], d["col_upper"],
                             d["row_lower"], d["row_upper"], d.get("sense", 1), d.get("offset", 0.0))


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_a_nan_warm_start_is_the_cold_start(gpu, precision):
    lp = gpu.Lp.generate(3, 21)
    m, n, _ = lp.dims
    cold = gpu.solve(lp, pdhg_tolerance=1e-4, pdhg_precision=precision, **REL)
    b = gpu.Batch(pdhg_tolerance=1e-4, pdhg_precision=precision, **REL)
    b.add(lp)
    b.warm_start(0, np.full(n, np.nan), np.full(m, np.nan))
    b.upload()
    b.solve()
    warm = b.result(0)
    assert warm.iterations == cold.iterations
    np.testing.assert_array_equal(warm.col_value, cold.col_value)
    np.testing.assert_array_equal(warm.row_dual, cold.row_dual)


def test_a_start_at_the_answer_stops_at_the_first_evaluation(gpu, oracle_mod):
    lp = gpu.Lp.generate(1, 1)
    d = lp.export()
    tight = gpu.solve(lp, pdhg_tolerance=1e-7, pdhg_precision="fp64", **REL)
    assert tight.converged
    # the autotuner rebuilds the batch's engines: the warm start must survive it
    b = gpu.Batch(pdhg_tolerance=1e-6, pdhg_precision="fp64", tile_nnz=-1, **REL)
    b.add(lp)
    b.warm_start(0, tight.col_value, tight.row_dual, tight.primal_weight, tight.step_size)
    b.upload()
    b.solve()
    r = b.result(0)
    assert r.converged and r.iterations <= 40, (r.iterations, r.info["message"])
    assert oracle_mod.evaluate(d, r.col_value, -r.row_dual)["worst"] <= 1e-6 * (1 + 1e-9)


def test_the_finish_of_the_2m_lp_reaches_its_known_optimum(gpu, oracle_mod, tmp_path):
    base = gpu.Lp.generate(5, 5).export()
    mult = lambda lp, v, t: oracle_mod.multiply(lp, v, transpose=t)  # noqa: E731
    d, optimum, _, _ = kkt_backwards_from(base, 55, mult)
    del base
    lp = to_lp(gpu, d)
    assert lp.dims == (2000000, 5000000, 25000000)
    first = gpu.solve(lp, pdhg_tolerance=1e-4, **REL)
    assert first.converged, first.info["message"]
    ev1 = oracle_mod.evaluate(d, first.col_value, -first.row_dual)
    assert ev1["worst"] <= 1e-4 * (1 + 1e-9)
    r = gpu.refine(lp, first, pdhg_tolerance=1e-7, **REL)
    ev = oracle_mod.evaluate(d, r.col_value, -r.row_dual)
    rel_obj = abs(r.objective - optimum) / abs(optimum)
    print(f"first: {first.iterations} iterations, {first.solve_seconds:.2f} s, objective error "
          f"{abs(first.objective - optimum) / abs(optimum):.2e}; finish: {r.polish_iterations} "
          f"FP64 iterations, {r.solve_seconds - first.solve_seconds:.2f} s, oracle KKT "
          f"{ev['worst']:.2e}, objective error {rel_obj:.2e}")
    assert r.converged and r.algorithm == "pdhg-gpu+fp64", r.info["message"]
    assert r.iterations == first.iterations + r.polish_iterations
    assert ev["worst"] <= 1e-7 * (1 + 1e-9)
    assert rel_obj <= 1e-6
    # the independent checker on what the answer claims (objective, activities, bounds)
    sol = tmp_path / "power_law_2m.sol"
    gpu.write_solution(sol, lp, r, model_name="power_law_2m")
    report = checker.verify(d, checker.read_sol(sol), primal_tol=1e-3)
    outcome = dict(report.outcome())
    assert outcome["structure"] and outcome["objective"] and outcome["activity agreement"], \
        [line for line in report.lines if not line[0]]
