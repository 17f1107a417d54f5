"""The GPU engine behind presolve (sk_pdhg_solve with presolve=true: solve.cpp:423-443's
sequence around the engine). The one-call path is bit for bit presolve -> the engine on the
reduced LP -> postsolve; on Netlib the FP64 engine behind presolve makes the oracle's decisions
(iterations inside the range the oracle gives under one-ulp cost moves, on the reduced LP) and
lands on the reference's published objective; on the 200,000 x 500,000 power-law LP, where
presolve fires 88 reductions (51 doubleton equations with fill-in), the FP32 answer postsolved to
the original LP meets 1e-4 there, as the oracle measures it."""
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
PUBLISHED = json.loads((GOLDEN / "pdhg_published.json").read_text())["instances"]
ULP_MOVES = (1.0, 1.0 + 2.0 ** -52, 1.0 - 2.0 ** -53, 1.0 + 2.0 ** -51, 1.0 - 2.0 ** -52,
             1.0 + 1e-13, 1.0 - 1e-13)
REL = dict(pdhg_stop_at_request=True, pdhg_absolute_primal=False)


@pytest.mark.parametrize("name", ["afiro", "sc50a", "blend"])
def test_the_one_call_path_is_presolve_engine_postsolve(gpu, name):
    lp = gpu.Lp.read_mps(GOLDEN / "netlib" / f"{name}.mps")
    one = gpu.solve(lp, presolve=True, pdhg_precision="fp64")
    p = gpu.presolve(lp)
    r = gpu.solve(p.reduced, pdhg_precision="fp64")
    x, y, d = p.postsolve(r.col_value, r.row_dual, r.col_dual)
    assert one.iterations == r.iterations and one.status == r.status
    np.testing.assert_array_equal(one.col_value, x)
    np.testing.assert_array_equal(one.row_dual, y)
    np.testing.assert_array_equal(one.col_dual, d)
    assert one.info["message"].startswith(f"presolve {lp.dims[0]} x {lp.dims[1]} -> ")


@pytest.mark.parametrize("limit", [200, 400])
@pytest.mark.parametrize("name", ["afiro", "sc50a", "sc50b", "blend", "sc105"])
def test_fp64_engine_behind_presolve_makes_the_oracles_decisions(gpu, oracle_mod, name, limit):
    """On the reduced LP, to an iteration limit: the engine's iterations, restarts and
    evaluations lie within the range the oracle gives when its costs move by one ulp to 1e-13
    (the whole runs are chaotic: a one-ulp move changes sc105's count from 61,440 to
    90,600-126,240, so only a bounded stretch of trajectory is comparable)."""
    lp = gpu.Lp.read_mps(GOLDEN / "netlib" / f"{name}.mps")
    red = gpu.presolve(lp).reduced.export()
    runs = [oracle_mod.solve_pdhg(dict(red, cost=red["cost"] * f), tolerance=1e-4,
                                  iteration_limit=limit) for f in ULP_MOVES]
    r = gpu.solve(lp, presolve=True, pdhg_precision="fp64", iteration_limit=limit)
    for k in ("iterations", "restarts", "evaluations"):
        lo, hi = min(x[k] for x in runs), max(x[k] for x in runs)
        assert lo <= getattr(r, k) <= hi, (k, getattr(r, k), lo, hi)


@pytest.mark.parametrize("name", ["afiro", "sc50a", "sc50b", "blend", "sc105"])
def test_fp64_engine_behind_presolve_reaches_the_published_answer(gpu, oracle_mod, name):
    lp = gpu.Lp.read_mps(GOLDEN / "netlib" / f"{name}.mps")
    d = lp.export()
    r = gpu.solve(lp, presolve=True, pdhg_precision="fp64")
    assert r.status == 1, r.info["message"]
    print(f"{name}: {r.iterations} iterations (the reference published "
          f"{PUBLISHED[name]['iterations']})")
    published = float(PUBLISHED[name]["objective"])
    assert abs(r.objective - published) <= 1e-7 * abs(published), (r.objective, published)
    # the reported quality is that of the postsolved answer on the original LP
    q = oracle_mod.recompute_quality(d, r.col_value, r.row_dual, r.col_dual, r.dual_bound)
    for f in ("objective", "primal_infeasibility", "dual_infeasibility", "complementarity_violation"):
        assert abs(getattr(r, f) - q[f]) <= 1e-12 * max(1.0, abs(q[f])), f


def test_power_law_200k_behind_presolve(gpu, oracle_mod):
    """88 reductions, 51 of them doubleton equations with fill-in. A doubleton fold moves the
    eliminated column's cost onto its partner scaled by the coefficient ratio (presolve.cpp:
    666-668), and this LP's coefficients span 1e-3..1e3: the reduced LP's costs reach ~1e6
    where the original's are within [-1, 1], and its relative KKT is measured against that
    scale. The engine meets 1e-4 on the reduced LP (the stop rule of solve_pdhg, which the
    reference also runs on the reduced LP); the postsolved point keeps the objective exactly
    (reduced objective + the folded constant), and its original-LP KKT is what the fold leaves
    (printed; FP64 iterates keep it near the target, FP32 ones do not)."""
    lp = gpu.Lp.generate(2, 11)
    d = lp.export()
    p = gpu.presolve(lp)
    assert p.info["by_kind"]["doubleton_equation"] == 51 and p.info["records"] == 88
    red = p.reduced.export()
    assert np.max(np.abs(red["cost"])) > 1e3 * np.max(np.abs(d["cost"]))
    plain = gpu.solve(lp, pdhg_tolerance=1e-4, **REL)
    for prec in ("fp32", "fp64"):
        r = gpu.solve(lp, presolve=True, pdhg_tolerance=1e-4, pdhg_precision=prec, **REL)
        assert r.converged, r.info["message"]
        inner = gpu.solve(p.reduced, pdhg_tolerance=1e-4, pdhg_precision=prec, **REL)
        assert inner.iterations == r.iterations
        evr = oracle_mod.evaluate(red, inner.col_value, -inner.row_dual)
        assert evr["worst"] <= 1e-4 * (1 + 1e-6)
        reduced_objective = red["offset"] + float(np.dot(red["cost"], inner.col_value))
        assert abs(r.objective - reduced_objective) <= 1e-9 * abs(reduced_objective)
        assert abs(r.objective - plain.objective) <= 1e-3 * abs(plain.objective)
        ev = oracle_mod.evaluate(d, r.col_value, -r.row_dual)
        print(f"{prec}: presolve {p.info['seconds']:.3f} s; {r.iterations} iterations (without "
              f"presolve {plain.iterations}); original-LP KKT primal {ev['primal']:.2e} dual "
              f"{ev['dual']:.2e} gap {ev['gap']:.2e}")


def test_stop_at_request_gives_the_cheap_answer_and_never_calls_it_optimal(gpu, oracle_mod):
    """test_pdhg.cpp:340-382 on the FP64 engine behind presolve (adlittle): the switch stops
    earlier than the run to the standard, reports `feasible`, not `optimal`, with a point that
    is primal feasible to 1e-7 and within 1e-3 of the honest objective, and says why."""
    lp = gpu.Lp.read_mps(GOLDEN / "netlib" / "adlittle.mps")
    opts = dict(presolve=True, pdhg_precision="fp64", pdhg_tolerance=1e-4, iteration_limit=1000000)
    honest = gpu.solve(lp, **opts)
    cheap = gpu.solve(lp, pdhg_stop_at_request=True, **opts)
    assert honest.status_name == "optimal", honest.info["message"]
    assert cheap.status_name == "feasible", cheap.info["message"]
    assert cheap.iterations < honest.iterations
    assert cheap.primal_infeasibility <= 1e-7, cheap.info["message"]
    assert abs(cheap.objective - honest.objective) <= 1e-3 * max(1.0, abs(honest.objective))
    assert "pdhg_stop_at_request" in cheap.info["message"]
    print(f"adlittle behind presolve: {cheap.iterations} iterations to the request, "
          f"{honest.iterations} to the standard (the reference: 132,520 and 192,080)")
