"""The interior-point finish on the GPU (sk_ipm_solve, sk_pdhg_polish): cold solves against
closed-form optima and the published Netlib optima, and PDHG answers polished to optimality
and judged by the independent answer checker (oracle/checker.py)."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import checker
from tests.mps_ref import read_mps
from tests.lp_cases import CLOSED_FORM, closed_form, dense_to_lp

pytestmark = pytest.mark.gpu
NETLIB = Path(__file__).resolve().parent / "golden" / "netlib"
OPTIMA = json.loads((NETLIB / "optima.json").read_text())["instances"]
FIRST = dict(pdhg_tolerance=1e-4, pdhg_stop_at_request=True, pdhg_absolute_primal=False)


def to_lp(sk, d):
    return sk.Lp.from_arrays(d["num_rows"], d["num_cols"], d["col_starts"], d["row_indices"],
                             d["values"], d["cost"], d["col_lower"], d["col_upper"],
                             d["row_lower"], d["row_upper"], d.get("sense", 1), d.get("offset", 0.0))


def rel(a, b):
    return abs(a - b) / max(1.0, abs(b))


@pytest.mark.parametrize("name", [c[0] for c in CLOSED_FORM])
def test_cold_interior_point_reaches_closed_form_optima(gpu, name):
    d, opt, xs = closed_form(name)
    r = gpu.solve_ipm(to_lp(gpu, d), iteration_limit=300)
    assert r.status_name == "optimal", r.info["message"]
    assert r.algorithm == "ipm"
    assert rel(r.objective, opt) <= 1e-7
    if xs is not None:
        np.testing.assert_allclose(r.col_value, xs, atol=1e-6)


@pytest.mark.parametrize("name", sorted(OPTIMA))
def test_cold_interior_point_reaches_published_netlib_optima(gpu, name):
    d = read_mps(NETLIB / f"{name}.mps")
    r = gpu.solve_ipm(to_lp(gpu, d), iteration_limit=300)
    assert r.status_name == "optimal", r.info["message"]
    # the published optima carry 11 significant digits
    assert rel(r.objective, OPTIMA[name]["published_optimal"]) <= 1e-8, r.objective


def polish_and_check(gpu, tmp_path, lp, d, name, col_names=None, row_names=None):
    first = gpu.solve(lp, **FIRST)
    assert first.status_name in ("feasible", "optimal"), first.info["message"]
    r = gpu.polish(lp, first)
    assert r.status_name == "optimal", r.info["message"]
    if first.status_name != "optimal":
        assert r.algorithm == "pdhg-gpu+ipm"
        assert r.polish_iterations > 0
        assert r.iterations == first.iterations + r.polish_iterations
        assert "polished by the interior point" in r.info["message"]
    sol = tmp_path / f"{name}.sol"
    gpu.write_solution(sol, lp, r, model_name=name, col_names=col_names, row_names=row_names)
    report = checker.verify(d, checker.read_sol(sol))
    assert report.verified, [line for line in report.lines if not line[0]]
    return first, r


@pytest.mark.parametrize("name", ["afiro", "sc50a", "sc50b", "blend", "sc105", "adlittle"])
def test_polished_netlib_answers_pass_the_checker(gpu, tmp_path, name):
    d = read_mps(NETLIB / f"{name}.mps")
    _, r = polish_and_check(gpu, tmp_path, to_lp(gpu, d), d, name, d["col_names"], d["row_names"])
    # the finish must reach 1e-6 relative objective against the published optimum
    assert rel(r.objective, OPTIMA[name]["published_optimal"]) <= 1e-6


@pytest.mark.parametrize("config, seed", [(1, 1), (3, 0), (3, 7)])
def test_polished_generated_answers_pass_the_checker(gpu, tmp_path, config, seed):
    lp = gpu.Lp.generate(config, seed)
    d = lp.export()
    first, r = polish_and_check(gpu, tmp_path, lp, d, f"c{config}_{seed}")
    # The first-order answer met relative gap <= 1e-4, |p - d| <= 1e-4 (1 + |p| + |d|), and
    # the optimum lies between its primal and dual objectives up to its (relative) primal and
    # dual violations, also <= 1e-4: so it is within 2e-4 (1 + |p| + |d|) of the optimum.
    bound = 2e-4 * (1 + abs(first.primal_objective) + abs(first.dual_objective))
    assert abs(first.objective - r.objective) <= bound


def test_polish_keeps_an_optimal_answer(gpu):
    d = read_mps(NETLIB / "afiro.mps")
    lp = to_lp(gpu, d)
    first = gpu.solve(lp, pdhg_tolerance=1e-9, pdhg_precision="fp64", iteration_limit=400000)
    assert first.status_name == "optimal"
    r = gpu.polish(lp, first)
    assert r.status_name == "optimal" and r.algorithm == first.algorithm
    assert r.iterations == first.iterations and r.polish_iterations == 0
    np.testing.assert_array_equal(r.col_value, first.col_value)


def test_polish_declines_a_factor_above_its_cap(gpu):
    lp = gpu.Lp.generate(1, 1)
    first = gpu.solve(lp, **FIRST)
    assert first.status_name == "feasible"
    r = gpu.polish(lp, first, max_factor_nonzeros=1000)
    assert r.status_name == "feasible" and r.algorithm == first.algorithm
    assert "declined" in r.info["message"]
    np.testing.assert_array_equal(r.col_value, first.col_value)


@pytest.mark.parametrize("name", ["israel", "share2b", "stocfor1"])
def test_library_mps_reader_feeds_the_published_optimum(gpu, name):
    # the C-ABI reader (sk_lp_read_mps) straight into the solvers: PDHG at 1e-4, then the
    # polish, lands on the published optimum
    lp = gpu.Lp.read_mps(NETLIB / f"{name}.mps")
    first = gpu.solve(lp, **FIRST)
    r = gpu.polish(lp, first)
    assert r.status_name == "optimal", r.info["message"]
    assert rel(r.objective, OPTIMA[name]["published_optimal"]) <= 1e-6


def test_every_answer_of_the_256_lp_batch_passes_the_checker_after_the_finish(gpu, tmp_path):
    """The headline workload end to end: the 256-LP batch solved together (FP32 split engine,
    1e-4 relative KKT), then each answer finished by the interior point and written as the
    reference's .sol file, and every one VERIFIED by the independent checker, with the
    finished objective within the first-order answer's own bound of it."""
    import time
    lps = [gpu.Lp.generate(3, k) for k in range(256)]
    b = gpu.Batch(**FIRST)
    for lp in lps:
        b.add(lp)
    b.upload()
    b.solve()
    t0, worst = time.perf_counter(), 0.0
    for k, lp in enumerate(lps):
        first = b.result(k)
        assert first.converged, (k, first.info["message"])
        r = gpu.polish(lp, first)
        assert r.status_name == "optimal", (k, r.info["message"])
        d = lp.export()
        sol = tmp_path / f"c3_{k}.sol"
        gpu.write_solution(sol, lp, r, model_name=f"c3_{k}")
        report = checker.verify(d, checker.read_sol(sol))
        assert report.verified, (k, [line for line in report.lines if not line[0]])
        bound = 2e-4 * (1 + abs(first.primal_objective) + abs(first.dual_objective))
        assert abs(first.objective - r.objective) <= bound, k
        worst = max(worst, abs(first.objective - r.objective) / max(1.0, abs(r.objective)))
        sol.unlink()
    print(f"256 answers finished and verified in {time.perf_counter() - t0:.1f} s; first-order "
          f"objectives within {worst:.2e} relative of the finished ones")


def _polish_lp(gpu):
    """test_pdhg.cpp:642-684: min -3x - 5y s.t. x <= 4, 2y <= 12, 3x + 2y <= 18, x, y >= 0;
    optimum -36 at (2, 6); PDHG is stopped at 20 iterations on purpose."""
    d = dense_to_lp([[1.0, 0.0], [0.0, 2.0], [3.0, 2.0]], [-np.inf] * 3, [4.0, 12.0, 18.0],
                    [-3.0, -5.0])
    return to_lp(gpu, d), d


def test_polish_finishes_an_iteration_limited_point_to_the_standard(gpu):
    lp, _ = _polish_lp(gpu)
    first = gpu.solve(lp, iteration_limit=20)
    assert first.status_name == "iteration_limit" and first.iterations == 20
    r = gpu.polish(lp, first)
    assert r.status_name == "optimal", r.info["message"]
    assert r.algorithm == "pdhg-gpu+ipm"
    assert abs(r.objective + 36.0) <= 1e-6
    np.testing.assert_allclose(r.col_value, [2.0, 6.0], atol=1e-5)
    assert r.polish_iterations > 0 and r.iterations == 20 + r.polish_iterations
    assert r.primal_infeasibility <= 1e-7 and r.dual_infeasibility_scaled <= 1e-7
    assert "polished by the interior point" in r.info["message"]


def test_without_the_polish_the_limited_point_is_reported_as_what_it_is(gpu):
    lp, _ = _polish_lp(gpu)
    rough = gpu.solve(lp, iteration_limit=20)
    assert rough.status_name == "iteration_limit"
    assert rough.algorithm == "pdhg-gpu" and rough.iterations == 20 and rough.polish_iterations == 0


def test_polish_declines_when_the_factor_cap_says_so_and_the_first_order_answer_stands(gpu):
    lp, _ = _polish_lp(gpu)
    first = gpu.solve(lp, iteration_limit=20)
    r = gpu.polish(lp, first, max_factor_nonzeros=0)
    assert r.status_name == "iteration_limit" and r.algorithm == "pdhg-gpu"
    assert r.iterations == 20 and "declined" in r.info["message"]
