"""The GPU restarted-PDHG engine against the FP64 CPU oracle (restatement of
src/pdhg/pdhg.cpp) on seeded inputs, and against closed-form optima."""
import numpy as np
import pytest

from tests.lp_cases import CLOSED_FORM, closed_form, kkt_lp

pytestmark = pytest.mark.gpu


def to_lp(sk, d):
    return sk.Lp.from_arrays(d["num_rows"], d["num_cols"], d["col_starts"], d["row_indices"],
                             d["values"], d["cost"], d["col_lower"], d["col_upper"],
                             d["row_lower"], d["row_upper"], d.get("sense", 1), d.get("offset", 0.0))


def kkt_of(oracle_mod, d, r):
    """Relative KKT (pdhg.cpp:110 worst()) of a returned point, measured by the oracle in FP64
    on the original model. row_dual = -sense * y (pdhg.cpp:657-661)."""
    y = -d.get("sense", 1) * r.row_dual
    return oracle_mod.evaluate(d, r.col_value, y)


@pytest.mark.parametrize("name", [c[0] for c in CLOSED_FORM])
def test_closed_form_fp64_reference_rule(gpu, name):
    # test_pdhg.cpp:95-146 with the reference's own stop rule at 1e-8
    d, optimum, xs = closed_form(name)
    r = gpu.solve(to_lp(gpu, d), pdhg_tolerance=1e-8, pdhg_precision="fp64", iteration_limit=200000)
    assert r.status_name == "optimal", r.info["message"]
    assert r.info["algorithm"] == "pdhg-gpu"
    assert abs(r.objective - optimum) <= 1e-6 * max(1.0, abs(optimum))
    if xs is not None:
        np.testing.assert_allclose(r.col_value, xs, atol=1e-4)
    assert r.primal_infeasibility <= 1e-7


@pytest.mark.parametrize("name", [c[0] for c in CLOSED_FORM])
def test_closed_form_fp32_relative_rule(gpu, oracle_mod, name):
    d, optimum, _ = closed_form(name)
    r = gpu.solve(to_lp(gpu, d), pdhg_tolerance=1e-6, pdhg_absolute_primal=False,
                  pdhg_stop_at_request=True, iteration_limit=200000)
    assert r.converged, r.info["message"]
    ev = kkt_of(oracle_mod, d, r)
    assert ev["worst"] <= 1e-6 * 1.000001
    assert abs(r.objective - optimum) <= 1e-4 * max(1.0, abs(optimum))


def test_iteration_limit_is_honest(gpu):
    # test_pdhg.cpp:148-157
    d, _, _ = closed_form("sum")
    r = gpu.solve(to_lp(gpu, d), pdhg_tolerance=1e-12, iteration_limit=20)
    assert r.status_name != "optimal"
    assert np.isinf(r.dual_bound)
    assert r.info["message"]
    assert r.iterations == 20


def test_step_size_survives_first_iteration(gpu):
    # test_pdhg.cpp:248-263
    d, _, _ = closed_form("sum")
    r = gpu.solve(to_lp(gpu, d), pdhg_tolerance=1e-8, pdhg_precision="fp64")
    assert r.status_name == "optimal", r.info["message"]
    assert r.iterations < 500


def test_large_rhs_never_claims_optimal_unverified(gpu):
    # test_pdhg.cpp:224-242, 265-294
    d = {"num_rows": 1, "num_cols": 2, "col_starts": np.array([0, 1, 2], np.int32),
         "row_indices": np.array([0, 0], np.int32), "values": np.array([1.0, 1.0]),
         "cost": np.array([1.0, 1.0]), "col_lower": np.zeros(2), "col_upper": np.full(2, np.inf),
         "row_lower": np.array([1e5]), "row_upper": np.array([np.inf])}
    for tol in (1e-4, 1e-6, 1e-8):
        r = gpu.solve(to_lp(gpu, d), pdhg_tolerance=tol, pdhg_precision="fp64", iteration_limit=200000)
        if r.status_name == "optimal":
            assert r.primal_infeasibility <= 1e-7
        assert r.status_name != "feasible" or r.primal_infeasibility <= 1e-7


def test_kkt_backwards_instances_fp64(gpu, oracle_mod):
    # the AgreesWithTheSimplexOnGeneratedInstances idea (test_pdhg.cpp:178-222): analytic optima
    rng = np.random.default_rng(20260904)
    compared = 0
    for _ in range(40):
        d, opt = kkt_lp(rng, int(rng.integers(2, 7)), int(rng.integers(2, 7)))
        r = gpu.solve(to_lp(gpu, d), pdhg_tolerance=1e-9, pdhg_precision="fp64",
                      iteration_limit=200000)
        if r.status_name != "optimal":
            continue
        compared += 1
        assert abs(r.objective - opt) <= 1e-6 * max(1.0, abs(opt)), (r.objective, opt)
    assert compared > 20


def test_fp64_engine_tracks_the_oracle_on_c1(gpu, oracle_mod):
    d = gpu.Lp.generate(1, 1).export()
    ref = oracle_mod.solve_pdhg(d, tolerance=1e-6, stop_at_request=True, absolute_primal=False)
    r = gpu.solve(gpu.Lp.generate(1, 1), pdhg_tolerance=1e-6, pdhg_stop_at_request=True,
                  pdhg_absolute_primal=False, pdhg_precision="fp64")
    assert r.converged and ref["converged"]
    # same method in FP64: only reduction order differs, so the spectral norm agrees closely
    assert abs(r.spectral_norm - ref["spectral_norm"]) <= 1e-12 * ref["spectral_norm"]
    assert abs(r.objective - ref["objective"]) <= 1e-5 * abs(ref["objective"])
    assert abs(r.iterations - ref["iterations"]) <= 0.25 * ref["iterations"] + 80
    ev = kkt_of(oracle_mod, d, r)
    assert ev["worst"] <= 1e-6 * 1.000001


def trajectory_case(gpu, name):
    if name == "c1":
        return gpu.Lp.generate(1, 1).export()
    if name.startswith("c3_"):
        return gpu.Lp.generate(3, int(name[3:])).export()
    return closed_form(name)[0]


# cost multipliers from one ulp to 1e-12 away from 1: the oracle run on each shows how far
# rounding-size perturbations move its own trajectory (counts and iterates)
ULP_MOVES = (1 + 2.0 ** -52, 1 - 2.0 ** -53, 1 + 2.0 ** -51, 1 - 2.0 ** -52, 1 + 1e-13,
             1 - 1e-13, 1 + 1e-12, 1 - 1e-12)


@pytest.mark.parametrize("limit", [100, 200, 400])
@pytest.mark.parametrize("name", ["c1", "c3_5"] + [c[0] for c in CLOSED_FORM])
def test_fp64_engine_follows_the_oracle_trajectory(gpu, oracle_mod, name, limit):
    """Same method, FP64 both sides, restarts on, the same iteration limit: the step-size
    rule (pdhg.cpp:454-496), the evaluation gate and the restart rule (:599-643) must make
    the oracle's decisions. A different reduction order is a rounding-size perturbation, which
    the method can amplify (a near-tie in an accept, a no-information step, a candidate
    choice), so the bar is the oracle's own behaviour when its costs move by one ulp to 1e-12:
    the engine's iterations, restarts and evaluations lie within the range those runs give
    (one value where the trajectory is stable), its status is theirs, and where their reported
    points agree to 1e-6 of scale the engine's is within 10x their spread (1e-9 where closer).
    Where rounding alone moves the oracle's point further (the 500 x 1000 LP after ~300
    iterations: ~1e-3), only the decisions are compared."""
    d = trajectory_case(gpu, name)
    ref = oracle_mod.solve_pdhg(d, tolerance=1e-12, iteration_limit=limit)
    runs = [ref] + [oracle_mod.solve_pdhg(dict(d, cost=d["cost"] * f), tolerance=1e-12,
                                          iteration_limit=limit) for f in ULP_MOVES]
    r = gpu.solve(to_lp(gpu, d), pdhg_tolerance=1e-12, iteration_limit=limit,
                  pdhg_precision="fp64")
    for k in ("iterations", "restarts", "evaluations"):
        lo, hi = min(x[k] for x in runs), max(x[k] for x in runs)
        print(f"{name} {limit} {k}: gpu {getattr(r, k)}, oracle {ref[k]} (range {lo}..{hi})")
        assert lo <= getattr(r, k) <= hi, (k, getattr(r, k), lo, hi)
    assert r.status in {x["status"] for x in runs}, r.info["message"]
    for k, got in (("col_value", r.col_value), ("row_dual", r.row_dual), ("col_dual", r.col_dual)):
        want = ref[k]
        if not want.size:
            continue
        spread = max(float(np.max(np.abs(x[k] - want))) for x in runs[1:])
        scale = max(1.0, float(np.max(np.abs(want))))
        diff = float(np.max(np.abs(got - want)))
        print(f"{name} {limit} {k}: |gpu - oracle| {diff:.3e}, oracle spread {spread:.3e}")
        if spread <= 1e-6 * scale:
            assert diff <= max(10.0 * spread, 1e-9 * scale), (k, diff, spread)
    if ref["status"] == 1:
        assert abs(r.objective - ref["objective"]) <= 1e-9 * max(1.0, abs(ref["objective"]))
    assert abs(r.spectral_norm - ref["spectral_norm"]) <= 1e-12 * ref["spectral_norm"]


@pytest.mark.parametrize("name", ["c1", "c3_5", "maximise", "offset"])
def test_quality_measures_are_the_oracle_recompute_quality(gpu, oracle_mod, name):
    """sk_pdhg_result's Solution quality measures (model.cpp:228-422) on the reported point
    are the oracle's recompute_quality of the same point and the same dual bound."""
    d = trajectory_case(gpu, name)
    r = gpu.solve(to_lp(gpu, d), pdhg_tolerance=1e-6, pdhg_stop_at_request=True,
                  pdhg_absolute_primal=False, pdhg_precision="fp64", iteration_limit=200000)
    q = oracle_mod.recompute_quality(d, r.col_value, r.row_dual, r.col_dual, r.dual_bound)
    for f in oracle_mod.QUALITY_FIELDS:
        assert r.info[f] == pytest.approx(q[f], rel=1e-12, abs=1e-300), f
    np.testing.assert_allclose(r.row_activity, q["row_activity"], rtol=1e-12, atol=1e-12)


def test_fp32_split_meets_1e4_on_c1(gpu, oracle_mod):
    d = gpu.Lp.generate(1, 1).export()
    tight = oracle_mod.solve_pdhg(d, tolerance=1e-8, stop_at_request=True, absolute_primal=False)
    r = gpu.solve(gpu.Lp.generate(1, 1), pdhg_tolerance=1e-4, pdhg_stop_at_request=True,
                  pdhg_absolute_primal=False)
    assert r.converged, r.info["message"]
    ev = kkt_of(oracle_mod, d, r)
    assert ev["worst"] <= 1e-4 * 1.000001
    # a 1e-4 relative KKT point is within a few 1e-4 of the optimum objective
    assert abs(r.objective - tight["objective"]) <= 1e-3 * abs(tight["objective"])


def test_batch_of_c3_instances_each_meet_1e4(gpu, oracle_mod):
    ks = list(range(12))
    lps = [gpu.Lp.generate(3, k) for k in ks]
    b = gpu.Batch(pdhg_tolerance=1e-4, pdhg_stop_at_request=True, pdhg_absolute_primal=False)
    for lp in lps:
        b.add(lp)
    b.upload()
    b.solve()
    for k, lp in zip(ks, lps):
        r = b.result(k)
        d = lp.export()
        assert r.converged, (k, r.info["message"])
        ev = kkt_of(oracle_mod, d, r)
        assert ev["worst"] <= 1e-4 * 1.000001, (k, ev["worst"])
    # solo equals in-batch, bit for bit (per-instance tiles and fixed reduction order)
    for k in (0, 5):
        solo = gpu.solve(lps[k], pdhg_tolerance=1e-4, pdhg_stop_at_request=True,
                         pdhg_absolute_primal=False)
        mixed = b.result(k)
        assert solo.iterations == mixed.iterations
        np.testing.assert_array_equal(solo.col_value, mixed.col_value)
        np.testing.assert_array_equal(solo.row_dual, mixed.row_dual)


def test_batch_resolve_is_reproducible(gpu):
    lps = [gpu.Lp.generate(3, k) for k in (20, 21, 22)]
    b = gpu.Batch(pdhg_tolerance=1e-4, pdhg_stop_at_request=True, pdhg_absolute_primal=False)
    for lp in lps:
        b.add(lp)
    b.upload()
    b.solve()
    first = [b.result(k).col_value.copy() for k in range(3)]
    b.solve()
    for k in range(3):
        np.testing.assert_array_equal(first[k], b.result(k).col_value)


def test_split_batch_matches_one_stream(gpu):
    """A batch run as three sub-batches on their own streams gives every instance bitwise the
    answer of the one-stream batch, and reports results in batch order."""
    lps = [gpu.Lp.generate(3, k) for k in range(30, 40)]
    kw = dict(pdhg_tolerance=1e-4, pdhg_stop_at_request=True, pdhg_absolute_primal=False)
    one = gpu.Batch(streams=1, **kw)
    split = gpu.Batch(streams=3, **kw)
    for b in (one, split):
        for lp in lps:
            b.add(lp)
        b.upload()
        b.solve()
    assert one.stats()["streams"] == 1 and split.stats()["streams"] == 3
    assert split.stats()["total_nnz"] == one.stats()["total_nnz"]
    for k in range(len(lps)):
        a, c = one.result(k), split.result(k)
        assert a.iterations == c.iterations and a.status == c.status
        np.testing.assert_array_equal(a.col_value, c.col_value)
        np.testing.assert_array_equal(a.row_dual, c.row_dual)
    # a split batch times its sub-batches' launches together; the bytes are the batch's
    ms_split, bytes_split = split.time_kernel(1, 4)
    ms_one, bytes_one = one.time_kernel(1, 4)
    # the same bytes, up to the one extra row pointer each sub-batch's CSR holds
    assert ms_split > 0 and ms_one > 0 and 0 <= bytes_split - bytes_one <= 4 * 2 * 2


def small_power_law(gpu, seed=5):
    """A config-5-shaped LP (power-law column degrees, |a| in [0.1, 10]) at 2,000 x 5,000."""
    return gpu.Lp.generate_params(**dict(gpu.Lp.preset(5, seed), num_rows=2000, num_cols=5000,
                                         num_equality_rows=1200, target_nnz=25000))


def test_relabelled_columns_return_the_lp_order(gpu, oracle_mod):
    """Columns with the most nonzeros first on the device (column_order 1) against the LP's
    own order (2): both answers, in the LP's order, meet 1e-4 relative KKT as the oracle
    measures them on the original model, and agree in objective. The power-iteration start is
    permuted with the columns, so the norm estimate agrees to rounding."""
    lp = small_power_law(gpu)
    d = lp.export()
    deg = np.diff(d["col_starts"])
    assert deg.max() > 4 * deg.mean()  # the relabelling is not the identity here
    kw = dict(pdhg_tolerance=1e-4, pdhg_stop_at_request=True, pdhg_absolute_primal=False)
    own = gpu.solve(lp, column_order=2, **kw)
    hot = gpu.solve(lp, column_order=1, **kw)
    for r in (own, hot):
        assert r.converged, r.info["message"]
        assert kkt_of(oracle_mod, d, r)["worst"] <= 1e-4 * 1.000001
    assert abs(hot.objective - own.objective) <= 1e-3 * max(1.0, abs(own.objective))
    own64 = gpu.solve(lp, column_order=2, pdhg_precision="fp64", iteration_limit=1)
    hot64 = gpu.solve(lp, column_order=1, pdhg_precision="fp64", iteration_limit=1)
    assert abs(hot64.spectral_norm - own64.spectral_norm) <= 1e-12 * own64.spectral_norm


def test_relabelled_instance_in_a_batch_equals_its_solo_run(gpu):
    kw = dict(pdhg_tolerance=1e-4, pdhg_stop_at_request=True, pdhg_absolute_primal=False,
              column_order=1)
    lps = [gpu.Lp.generate(3, 7), small_power_law(gpu, 6), gpu.Lp.generate(3, 8)]
    b = gpu.Batch(**kw)
    for lp in lps:
        b.add(lp)
    b.upload()
    b.solve()
    for k in range(3):
        solo = gpu.solve(lps[k], **kw)
        mixed = b.result(k)
        assert solo.converged and solo.iterations == mixed.iterations
        np.testing.assert_array_equal(solo.col_value, mixed.col_value)
        np.testing.assert_array_equal(solo.col_dual, mixed.col_dual)
        np.testing.assert_array_equal(solo.row_dual, mixed.row_dual)


def test_autotuned_batch_runs_with_the_capacity_it_picked(gpu, oracle_mod):
    """tile_nnz = -1: the upload picks one of the candidate capacities by timing; the batch
    then runs exactly as one uploaded with that capacity, and every answer meets 1e-4."""
    lps = [gpu.Lp.generate(3, k) for k in range(40, 56)]
    kw = dict(pdhg_tolerance=1e-4, pdhg_stop_at_request=True, pdhg_absolute_primal=False)
    tuned = gpu.Batch(tile_nnz=-1, **kw)
    for lp in lps:
        tuned.add(lp)
    tuned.upload()
    tuned.solve()
    cap = tuned.stats()["tile_nnz"]
    assert cap in (1536, 2048, 2560, 3072, 4096)
    fixed = gpu.Batch(tile_nnz=cap, **kw)
    for lp in lps:
        fixed.add(lp)
    fixed.upload()
    fixed.solve()
    assert fixed.stats()["tile_nnz"] == cap
    for k, lp in enumerate(lps):
        a, c = tuned.result(k), fixed.result(k)
        assert a.converged and a.iterations == c.iterations
        np.testing.assert_array_equal(a.col_value, c.col_value)
        assert kkt_of(oracle_mod, lp.export(), a)["worst"] <= 1e-4 * 1.000001


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_resident_solve_equals_separate_launches(gpu, precision):
    """Instances with a few tiles run as the resident solve (many steps and evaluations per
    launch, with the per-tile code of the separate step kernels): the answers are bit for bit
    those of the separate launches (with the decision folded into the next primal launch,
    step_fusion = 1, or in a launch of its own every step, 3), alone and as a batch of small
    instances. The capacity is pinned to 2,560, at which the 500 x 1,000 LPs have four tiles (at
    the default one-round capacity of 1,536 they have eight and run as separate launches)."""
    kw = dict(pdhg_tolerance=1e-4, pdhg_stop_at_request=True, pdhg_absolute_primal=False,
              pdhg_precision=precision, tile_nnz=2560)
    groups = [[gpu.Lp.generate(1, 1)],
              [gpu.Lp.generate(1, s) for s in (2, 3, 4)],
              [to_lp(gpu, closed_form(name)[0]) for name in ("sum", "range", "maximise")]]
    for lps in groups:
        runs = {}
        for fusion in (0, 1, 3):
            b = gpu.Batch(step_fusion=fusion, **kw)
            for lp in lps:
                b.add(lp)
            b.upload()
            b.solve()
            runs[fusion] = (b.stats()["fused"], [b.result(k) for k in range(len(lps))])
        assert runs[0][0] == 1 and runs[1][0] == 0 and runs[3][0] == 0
        for other in (1, 3):
            for fused, plain in zip(runs[0][1], runs[other][1]):
                assert fused.converged and fused.iterations == plain.iterations
                assert fused.status == plain.status and fused.attempts == plain.attempts
                np.testing.assert_array_equal(fused.col_value, plain.col_value)
                np.testing.assert_array_equal(fused.row_dual, plain.row_dual)
                np.testing.assert_array_equal(fused.col_dual, plain.col_dual)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("block,stride", [(40, 20), (8, 3)])
def test_decisions_in_the_primal_launch_equal_their_own_launch(gpu, precision, block, stride):
    """Most steps are decided by the next step's primal launch: every column chunk of an
    instance decides from the same records and partials, and the first chunk records the
    decision in the other buffer of the record pair. Iterations, attempts (rejected steps
    included, which reload the iterate the speculative loads skipped), restarts, statuses
    and answers are bit for bit those of deciding every step in a launch of its own
    (step_fusion = 2), for instances of many column chunks. The 8-step block with an
    evaluation every 3 steps has an odd number of such decisions, which the schedule evens
    out by deciding one more step in its own launch."""
    kw = dict(pdhg_tolerance=1e-4, pdhg_stop_at_request=True, pdhg_absolute_primal=False,
              pdhg_precision=precision, steps_per_block=block, eval_stride=stride)
    lps = [gpu.Lp.generate(3, k) for k in range(6)] + [gpu.Lp.generate(1, 1)]
    runs = {}
    for fusion in (0, 2):
        b = gpu.Batch(step_fusion=fusion, **kw)
        for lp in lps:
            b.add(lp)
        b.upload()
        b.solve()
        assert b.stats()["fused"] == 0
        runs[fusion] = [b.result(k) for k in range(len(lps))]
    rejected = 0
    for folded, alone in zip(runs[0], runs[2]):
        assert folded.converged and folded.iterations == alone.iterations
        assert folded.attempts == alone.attempts and folded.restarts == alone.restarts
        assert folded.status == alone.status and folded.evaluations == alone.evaluations
        np.testing.assert_array_equal(folded.col_value, alone.col_value)
        np.testing.assert_array_equal(folded.row_dual, alone.row_dual)
        np.testing.assert_array_equal(folded.col_dual, alone.col_dual)
        rejected += folded.attempts - folded.iterations
    assert rejected > 0, "no rejected step: the reload path was not exercised"
