"""Presolve and postsolve (sih26_amd/csrc/presolve.cpp, restating src/presolve/presolve.cpp),
on the CPU. The gate is the reference's own (tests/unit/test_presolve.cpp:63-81): the same LP
solved with presolve on and off must agree on the status and the objective, and the answer
postsolved to the original LP must be primal AND dual feasible there, with no gap. The LPs are
solved by an independent solver (HiGHS's dual simplex, tests/highs_lp.py), so postsolve is fed
the vertex answers the reference's simplex gives it; the quality of the postsolved answer is
measured on the original LP by the oracle's recompute_quality (model.cpp:228-422). Each
hand-written case of test_presolve.cpp:85-362 is restated with its closed-form values; the
reference's fuzz of free column singletons and doubleton equations (:367-486) is restated with
the same generator shape; and the nine Netlib fixtures round-trip to their published optima."""
import json
from pathlib import Path

import numpy as np
import pytest

pytest.importorskip("scipy.optimize")

import sih26_amd as sk  # noqa: E402
from oracle import oracle  # noqa: E402
from tests import highs_lp  # noqa: E402
from tests.lp_cases import dense_to_lp  # noqa: E402

INF = np.inf
NETLIB = Path(__file__).resolve().parent / "golden" / "netlib"


def to_lp(d):
    return sk.Lp.from_arrays(d["num_rows"], d["num_cols"], d["col_starts"], d["row_indices"],
                             d["values"], d["cost"], d["col_lower"], d["col_upper"],
                             d["row_lower"], d["row_upper"], d.get("sense", 1), d.get("offset", 0.0))


def make_lp(rows, row_lower, row_upper, cost, col_lower, col_upper):
    """test_presolve.cpp:40-61: an LP from dense rows."""
    return dense_to_lp(rows, row_lower, row_upper, cost, col_upper, col_lower=col_lower)


def with_presolve(d, integer=None, reduced_answer=None):
    """presolve -> HiGHS on the reduced LP -> postsolve, measured on the original LP."""
    p = sk.presolve(to_lp(d), integer)
    if p.proved_infeasible:
        return {"status": "infeasible", "presolved": p}
    red = p.reduced.export()
    ans = highs_lp.solve(red)
    if ans["status"] != "optimal":
        return {"status": ans["status"], "presolved": p}
    if reduced_answer is not None:
        ans = reduced_answer(red, ans)
    x, y, dd = p.postsolve(ans["x"], ans["row_dual"], ans["col_dual"])
    q = oracle.recompute_quality(d, x, y, dd, ans["objective"])
    return {"status": "optimal", "x": x, "row_dual": y, "col_dual": dd, "q": q,
            "objective": q["objective"], "presolved": p, "reduced": red}


def expect_agrees_with_unpresolved(d, integer=None):
    """test_presolve.cpp:63-81."""
    off = highs_lp.solve(d)
    on = with_presolve(d, integer)
    assert on["status"] == off["status"], (on["status"], off["status"])
    if off["status"] != "optimal":
        return on
    q = on["q"]
    assert abs(on["objective"] - off["objective"]) <= 1e-9 * max(1.0, abs(off["objective"]))
    assert q["primal_infeasibility"] <= 1e-7
    assert q["dual_infeasibility"] <= 1e-7
    assert abs(q["absolute_gap"]) <= 1e-7
    return on


def test_fixed_column_is_folded_into_the_objective_constant():
    d = make_lp([[1.0, 1.0]], [5.0], [INF], [4.0, 2.0], [0.0, 3.0], [INF, 3.0])
    on = expect_agrees_with_unpresolved(d)
    assert on["objective"] == pytest.approx(14.0, abs=1e-9)
    assert on["x"][1] == pytest.approx(3.0, abs=1e-9)
    # the whole LP goes: x1 fixed (its cost 2 * 3 into the constant), the row then a bound on x0
    kinds = [r["kind"] for r in on["presolved"].records()]
    assert kinds == ["fixed_column", "singleton_row", "empty_column"]
    assert on["reduced"]["num_cols"] == 0 and on["reduced"]["offset"] == 14.0


def test_singleton_row_becomes_a_bound_and_keeps_its_dual():
    d = make_lp([[2.0]], [6.0], [INF], [5.0], [0.0], [INF])
    on = expect_agrees_with_unpresolved(d)
    assert on["objective"] == pytest.approx(15.0, abs=1e-9)
    assert on["row_dual"][0] == pytest.approx(2.5, abs=1e-9)  # the removed row is priced
    assert on["q"]["dual_infeasibility"] <= 1e-7


def test_singleton_row_is_priced_from_a_point_that_only_nearly_touches_it():
    """test_presolve.cpp:114-135: an interior-point answer lands within its relative tolerance
    of a bound, not on it; the singleton row -7 x0 >= -4900 must still be priced (dual 1). The
    reduced answer is moved off the vertex by 1e-10 relative, as an interior point leaves it."""
    d = make_lp([[-7.0, 0.0], [1.0, 1.0]], [-4900.0, -INF], [INF, 1000.0], [-8.0, -1.0],
                [0.0, 0.0], [10000.0, 10000.0])

    def nearly(red, ans):
        x = ans["x"].copy()
        x *= 1.0 - 1e-10
        return dict(ans, x=x)

    on = with_presolve(d, reduced_answer=nearly)
    assert on["status"] == "optimal"
    assert on["objective"] == pytest.approx(-5900.0, abs=1e-5)
    assert on["row_dual"][0] == pytest.approx(1.0, abs=1e-6)
    assert on["q"]["dual_infeasibility_scaled"] <= 1e-7
    expect_agrees_with_unpresolved(d)


def test_two_opposing_singleton_rows_price_the_admissible_one():
    d = make_lp([[0.0, 6.0], [0.0, -3.0], [-3.0, -6.0]], [24.0, -12.0, -24.0], [INF, INF, INF],
                [-5.0, 6.0], [0.0, 0.0], [INF, INF])
    on = expect_agrees_with_unpresolved(d)
    assert on["objective"] == pytest.approx(24.0, abs=1e-9)
    assert on["q"]["dual_infeasibility"] <= 1e-7


def test_redundant_row_is_removed_and_priced_at_zero():
    d = make_lp([[1.0, 1.0], [1.0, 0.0]], [-5.0, 2.0], [INF, INF], [1.0, 1.0], [0.0, 0.0], [INF, INF])
    on = expect_agrees_with_unpresolved(d)
    assert on["row_dual"][0] == 0.0
    assert on["presolved"].records()[0]["kind"] == "redundant_row"


def test_empty_row_excluding_zero_is_infeasible():
    d = make_lp([[0.0]], [2.0], [INF], [1.0], [0.0], [INF])
    p = sk.presolve(to_lp(d))
    assert p.proved_infeasible
    assert p.message == "row 0 has no entries but requires activity in [2, inf], which excludes zero"


def test_an_unbounded_objective_is_not_reported_as_infeasible():
    """test_presolve.cpp:175-193: presolve cannot conclude unboundedness; the empty column of
    negative cost stays in the reduced LP for the engine to judge."""
    unbounded = make_lp([[0.0, 1.0]], [1.0], [INF], [-1.0, 1.0], [0.0, 0.0], [INF, INF])
    on = with_presolve(unbounded)
    assert on["status"] == "unbounded"
    assert on["presolved"].info["reduced_cols"] >= 1
    infeasible = make_lp([[0.0, -1.0]], [2.0], [INF], [-1.0, 1.0], [0.0, 0.0], [INF, 0.0])
    assert with_presolve(infeasible)["status"] == "infeasible"
    assert highs_lp.solve(infeasible)["status"] == "infeasible"


def test_integer_bounds_are_rounded_inward_never_outward():
    d = make_lp([[2.0]], [-INF], [7.0], [-1.0], [0.0], [10.0])
    p = sk.presolve(to_lp(d), integer_columns=[1])
    red = p.reduced.export()
    # 2 x0 <= 7 is x0 <= 3.5, and for an integer column x0 <= 3; the column then goes to its
    # bound 3 (an empty column of negative cost), so the LP relaxation left is the MIP's answer
    x, _, _ = p.postsolve(np.zeros(red["num_cols"]), np.zeros(red["num_rows"]), np.zeros(red["num_cols"]))
    assert x[0] == 3.0
    assert red["offset"] == -3.0
    continuous = sk.presolve(to_lp(d))
    xc, _, _ = continuous.postsolve(np.zeros(0), np.zeros(0), np.zeros(0))
    assert xc[0] == 3.5


def test_chained_reductions_still_round_trip():
    d = make_lp([[1.0, 1.0, 0.0], [0.0, 1.0, 0.0], [1.0, 0.0, 1.0]], [4.0, 2.0, 3.0],
                [INF, 2.0, INF], [1.0, 3.0, 2.0], [0.0, 0.0, 0.0], [INF, INF, INF])
    expect_agrees_with_unpresolved(d)


def test_free_column_singleton_is_substituted_and_removed():
    d = make_lp([[1.0, 1.0]], [2.0], [INF], [5.0, 2.0], [0.0, -INF], [10.0, INF])
    on = expect_agrees_with_unpresolved(d)
    assert on["objective"] == pytest.approx(4.0, abs=1e-9)
    np.testing.assert_allclose(on["x"], [0.0, 2.0], atol=1e-9)
    assert on["presolved"].records()[0]["kind"] == "free_column_singleton"


def test_doubleton_equation_eliminates_a_column():
    d = make_lp([[1.0, -2.0]], [0.0], [0.0], [1.0, 2.0], [0.0, -5.0], [10.0, 5.0])
    on = expect_agrees_with_unpresolved(d)
    assert on["objective"] == pytest.approx(0.0, abs=1e-9)
    np.testing.assert_allclose(on["x"], [0.0, 0.0], atol=1e-9)
    rec, park = on["presolved"].records()
    assert rec["kind"] == "doubleton_equation" and rec["column"] == 0 and rec["partner_column"] == 1
    assert rec["eliminated_cost"] == 1.0 and rec["substituted_rhs"] == 0.0
    # x1's bounds become [0, 5] (x0 = 2 x1 in [0, 10]) and its folded cost 2 - 1 (-2 / 1) = 4,
    # so with no row left it is an empty column parked at its new lower bound
    assert park["kind"] == "empty_column" and park["index"] == 1 and park["value"] == 0.0
    assert on["reduced"]["num_cols"] == 0


def test_doubleton_equation_cascades_into_a_singleton_row():
    d = make_lp([[1.0, 1.0], [1.0, -2.0]], [1.0, 0.0], [4.0, 0.0], [1.0, 2.0], [0.0, -INF], [10.0, INF])
    on = expect_agrees_with_unpresolved(d)
    assert on["objective"] == pytest.approx(4.0 / 3.0, abs=1e-9)
    np.testing.assert_allclose(on["x"], [2.0 / 3.0, 1.0 / 3.0], atol=1e-9)
    assert on["q"]["dual_infeasibility"] <= 1e-7


def test_fixed_column_sharing_a_folded_doubleton_row_still_prices_correctly():
    d = make_lp([[2.0, 3.0, 4.0]], [20.0], [20.0], [1.0, 2.0, 3.0], [5.0, 0.0, 0.0], [5.0, 10.0, 10.0])
    on = expect_agrees_with_unpresolved(d)
    assert on["objective"] == pytest.approx(35.0 / 3.0, abs=1e-9)
    np.testing.assert_allclose(on["x"], [5.0, 10.0 / 3.0, 0.0], atol=1e-9)
    assert on["q"]["dual_infeasibility"] <= 1e-7


def test_singleton_row_chain_ending_in_a_doubleton_prices_every_link():
    """test_presolve.cpp:304-353 (Netlib ganges at unit scale): the dual passes must run to a
    fixed point for the doubleton's price to reach every link of the chain."""
    d = make_lp([[1.0, 0.0, 0.0, 0.0, 0.0, 0.0],
                 [0.0, 1.0, 0.0, 0.0, 0.0, 0.0],
                 [-1.0, -1.0, 1.0, 0.0, 0.0, 0.0],
                 [0.0, 0.0, -1.0, 1.0, 1.0, 0.0],
                 [0.0, 0.0, 0.0, 0.0, 1.0, 1.0]],
                [160.0, 160.0, 0.0, 0.0, -INF], [160.0, 160.0, 0.0, 0.0, 500.0],
                [0.0, 0.0, 0.0, 1.0, 2.0, -1.0], [0.0] * 6, [INF, INF, INF, INF, INF, 100.0])
    on = expect_agrees_with_unpresolved(d)
    assert on["objective"] == pytest.approx(220.0, abs=1e-9)
    np.testing.assert_allclose(on["x"][[0, 2, 3]], [160.0, 320.0, 320.0], atol=1e-9)
    np.testing.assert_allclose(on["col_dual"][:4], 0.0, atol=1e-9)
    np.testing.assert_allclose(on["row_dual"][:4], 1.0, atol=1e-9)
    assert abs(on["row_dual"][4]) <= 1e-9
    assert on["q"]["complementarity_violation"] <= 1e-9
    assert on["q"]["dual_infeasibility_scaled"] <= 1e-7


def test_leaves_a_model_with_nothing_to_remove_alone():
    d = make_lp([[1.0, 2.0], [3.0, 1.0]], [4.0, 5.0], [INF, INF], [1.0, 1.0], [0.0, 0.0], [INF, INF])
    on = expect_agrees_with_unpresolved(d)
    p = on["presolved"]
    assert p.info["records"] == 0 and p.info["passes"] == 1
    red = on["reduced"]
    for k in ("col_starts", "row_indices", "values", "cost", "col_lower", "col_upper",
              "row_lower", "row_upper"):
        np.testing.assert_array_equal(red[k], d[k])


def random_doubleton_prone_lp(rng: np.random.Generator, trial: int) -> dict:
    """test_presolve.cpp:367-452: an LP around a known feasible point, some rows forced to be
    two-entry equalities tight at that point and some columns forced free."""
    n, m = 3 + trial % 5, 2 + trial % 4
    coefficient = lambda: rng.uniform(-4.0, 4.0)  # noqa: E731
    unit = lambda: rng.uniform(0.0, 1.0)  # noqa: E731
    cost = np.array([coefficient() for _ in range(n)])
    lo, hi, x0 = np.zeros(n), np.zeros(n), np.zeros(n)
    for j in range(n):
        if unit() < 0.35:
            lo[j], hi[j], x0[j] = -INF, INF, coefficient()
        else:
            centre, half = coefficient(), 1.0 + 4.0 * unit()
            lo[j], hi[j] = centre - half, centre + half
            x0[j] = lo[j] + unit() * (hi[j] - lo[j])
    rows = np.zeros((m, n))
    rl, ru = np.zeros(m), np.zeros(m)
    for i in range(m):
        if unit() < 0.5 and n >= 2:
            j1 = int(unit() * n) % n
            j2 = int(unit() * n) % n
            if j2 == j1:
                j2 = (j1 + 1) % n
            a1, a2 = coefficient(), coefficient()
            if abs(a1) < 0.5:
                a1 = -0.5 if a1 < 0.0 else 0.5
            if abs(a2) < 0.5:
                a2 = -0.5 if a2 < 0.0 else 0.5
            rows[i, j1], rows[i, j2] = a1, a2
            rl[i] = ru[i] = a1 * x0[j1] + a2 * x0[j2]
        else:
            for j in range(n):
                if unit() < 0.6:
                    rows[i, j] = coefficient()
            if not rows[i].any():
                rows[i, 0] = 1.0
            act = float(rows[i] @ x0)
            rl[i] = act - (0.5 + 3.0 * unit())
            ru[i] = act + (0.5 + 3.0 * unit())
    return make_lp(rows, rl, ru, cost, lo, hi)


def test_fuzz_free_column_singleton_and_doubleton_equation():
    """test_presolve.cpp:454-486: 400 trials; on and off agree on the status and, where
    optimal, on the objective (1e-6 relative) and on primal and dual feasibility."""
    rng = np.random.default_rng(923004)
    compared, fired = 0, {"free_column_singleton": 0, "doubleton_equation": 0}
    for trial in range(400):
        d = random_doubleton_prone_lp(rng, trial)
        off = highs_lp.solve(d)
        on = with_presolve(d)
        assert on["status"] == off["status"], (trial, on["status"], off["status"])
        for r in on["presolved"].records():
            if r["kind"] in fired:
                fired[r["kind"]] += 1
        if off["status"] != "optimal":
            continue
        compared += 1
        assert abs(on["objective"] - off["objective"]) <= 1e-6 * max(1.0, abs(off["objective"])), trial
        assert on["q"]["primal_infeasibility"] <= 1e-7, trial
        assert on["q"]["dual_infeasibility"] <= 1e-7, (trial, on["q"])
    assert compared > 250, compared
    assert fired["free_column_singleton"] > 50 and fired["doubleton_equation"] > 50, fired


OPTIMA = json.loads((NETLIB / "optima.json").read_text())["instances"] if NETLIB.exists() else {}


@pytest.mark.parametrize("name", sorted(OPTIMA))
def test_netlib_round_trips_to_the_published_optimum(name):
    lp = sk.Lp.read_mps(NETLIB / f"{name}.mps")
    d = lp.export()
    on = with_presolve(d)
    assert on["status"] == "optimal"
    published = OPTIMA[name]["published_optimal"]
    assert abs(on["objective"] - published) <= 1e-9 * max(1.0, abs(published)) + 5e-9
    q = on["q"]
    assert q["primal_infeasibility_scaled"] <= 1e-7, q
    assert q["dual_infeasibility_scaled"] <= 1e-7, q
    assert abs(q["relative_gap"]) <= 1e-9, q
    off = highs_lp.solve(d)
    assert abs(on["objective"] - off["objective"]) <= 1e-9 * max(1.0, abs(off["objective"]))


def test_a_solve_with_presolve_proves_infeasibility_without_an_engine():
    """solve.cpp:423-442: presolve's proof ends the solve (status infeasible, algorithm
    "presolve", no certificate); no device is touched, so this runs on the CPU."""
    d = make_lp([[0.0, -1.0]], [2.0], [INF], [-1.0, 1.0], [0.0, 0.0], [INF, 0.0])
    r = sk.solve(to_lp(d), presolve=True)
    assert r.status == 2 and r.info["algorithm"] == "presolve"
    assert r.info["message"].startswith("row 0 has no entries but requires activity in [2, inf]")
    assert "no Farkas certificate" in r.info["message"]


def test_a_solve_that_presolve_finishes_needs_no_engine():
    """Every row and column removed: the answer is postsolve's alone, measured on the original
    LP (objective 14 = 4 * 2 + 2 * 3), with the row priced (dual 4)."""
    d = make_lp([[1.0, 1.0]], [5.0], [INF], [4.0, 2.0], [0.0, 3.0], [INF, 3.0])
    r = sk.solve(to_lp(d), presolve=True)
    assert r.status == 1 and r.iterations == 0, r.info["message"]
    assert r.objective == 14.0
    np.testing.assert_array_equal(r.col_value, [2.0, 3.0])
    np.testing.assert_array_equal(r.row_dual, [4.0])
    np.testing.assert_array_equal(r.row_activity, [5.0])
    assert r.info["message"].startswith("presolve 1 x 2 -> 0 x 0; ")
    assert r.dual_infeasibility == 0.0 and r.primal_infeasibility == 0.0
