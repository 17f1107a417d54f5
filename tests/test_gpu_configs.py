"""The seeded workloads at their full sizes, each answer measured by the FP64 oracle on the
original model (pdhg.cpp:143-263 restated):
  - the 200,000 x 500,000 power-law LP (2.5M nonzeros, |a| log-uniform on [1e-3, 1e3], seed
    11): the FP32/FP64 split engine and the plain FP64 engine at 1e-4 relative KKT, the FP64
    engine at 1e-6, and its device scaling and step products at tile capacity 1,536;
  - the 2,000,000 x 5,000,000 power-law LP (25M nonzeros, seed 5) at 1e-4, FP32 split;
  - serving-stream requests (rows log-uniform on [500, 50,000]) through the server, each the
    same bits as its solo solve.
Stop rule: relative KKT (primal, dual, gap) at the tolerance, FP64, on the original model."""
import time

import numpy as np
import pytest

from tests.spmv_order import error_bound, step_product

pytestmark = pytest.mark.gpu

REL = dict(pdhg_stop_at_request=True, pdhg_absolute_primal=False)


def oracle_kkt(oracle_mod, d, r):
    return oracle_mod.evaluate(d, r.col_value, -d.get("sense", 1) * r.row_dual)


def check_answer(oracle_mod, d, r, tol):
    assert r.converged, r.info["message"]
    ev = oracle_kkt(oracle_mod, d, r)
    print(f"  iterations {r.iterations} restarts {r.restarts} seconds {r.solve_seconds:.2f}: "
          f"oracle primal {ev['primal']:.3e} dual {ev['dual']:.3e} gap {ev['gap']:.3e}")
    assert ev["worst"] <= tol * (1 + 1e-9), ev
    # the engine's own FP64 evaluation on the original model is the oracle's, to rounding
    for mine, theirs in ((r.info["rel_primal"], ev["primal"]), (r.info["rel_dual"], ev["dual"]),
                         (r.info["rel_gap"], ev["gap"])):
        assert abs(mine - theirs) <= 1e-7 * theirs + 1e-12, (mine, theirs)
    return ev


@pytest.fixture(scope="module")
def lp_200k(gpu):
    lp = gpu.Lp.generate(2, 11)
    assert lp.dims == (200000, 500000, 2500000)
    return lp, lp.export()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_power_law_200k_meets_1e4(gpu, oracle_mod, lp_200k, precision):
    lp, d = lp_200k
    r = gpu.solve(lp, pdhg_tolerance=1e-4, pdhg_precision=precision, **REL)
    check_answer(oracle_mod, d, r, 1e-4)


def test_power_law_200k_fp64_meets_1e6(gpu, oracle_mod, lp_200k):
    lp, d = lp_200k
    r = gpu.solve(lp, pdhg_tolerance=1e-6, pdhg_precision="fp64", **REL)
    check_answer(oracle_mod, d, r, 1e-6)


def test_power_law_200k_scaling_and_products_at_capacity_1536(gpu, oracle_mod, lp_200k):
    """Log-uniform values over six decades: the device scaling is the oracle's bit for bit,
    and the step products at the capacity the autotuner picks for this LP (vector tiles for
    columns above 1,536 nonzeros) are the fixed-order sums of those values."""
    lp, d = lp_200k
    b = gpu.Batch(iteration_limit=3, tile_nnz=1536)
    b.add(lp)
    b.upload()
    b.solve()
    s = oracle_mod.build_scaling(d)
    got = b.probe_scaling(0)
    for k in ("row", "column", "values"):
        np.testing.assert_array_equal(got[k], s[k], err_msg=k)
    assert np.diff(d["col_starts"]).max() > 1536
    rng = np.random.default_rng(21)
    m, n = d["num_rows"], d["num_cols"]

    def mult(vals, vec, transpose):
        return oracle_mod.multiply(dict(d, values=vals), vec, transpose=transpose)

    for which, vec in ((1, rng.standard_normal(n)), (2, rng.standard_normal(m))):
        out = b.probe_spmv(0, which, vec)
        want = step_product(d, s["values"], vec, which, np.float32).astype(np.float64)
        np.testing.assert_array_equal(out, want)
        exact = mult(s["values"], vec, which == 2)
        assert np.all(np.abs(out - exact) <= error_bound(d, s["values"], vec, which, 2.0 ** -24, mult))


def test_power_law_2m_fp32_meets_1e4(gpu, oracle_mod):
    lp = gpu.Lp.generate(5, 5)
    assert lp.dims == (2000000, 5000000, 25000000)
    d = lp.export()
    r = gpu.solve(lp, pdhg_tolerance=1e-4, **REL)
    check_answer(oracle_mod, d, r, 1e-4)


def _drain(srv, expected, timeout=300.0):
    out = {}
    t_end = time.time() + timeout
    while len(out) < len(expected) and time.time() < t_end:
        for rid in srv.completed():
            out[rid] = srv.result(rid, expected[rid])
        time.sleep(0.002)
    assert len(out) == len(expected), f"only {len(out)} of {len(expected)} completed"
    return out


def test_serving_stream_requests_equal_their_solo_solves(gpu, oracle_mod):
    """The first 20 requests of the serving stream (sizes up to 50,000 x 100,000), arriving
    5 ms apart into the default two slots: every answer is its solo solve bit for bit and
    meets 1e-4 on the oracle's measure."""
    srv = gpu.Server(pdhg_tolerance=1e-4, **REL)
    dims, exported = {}, {}
    t0 = srv.now()
    for i in range(20):
        lp = gpu.Lp.generate(4, i)
        m, n, _ = lp.dims
        exported[i] = lp.export()
        rid = srv.submit(lp, arrival=t0 + 0.005 * i)
        dims[rid] = (m, n)
        time.sleep(0.005)
    got = _drain(srv, dims)
    st = srv.stats()
    srv.close()
    assert st["completed"] == 20 and st["failed"] == 0
    assert max(exported[i]["num_rows"] for i in range(20)) > 20000
    for i, rid in enumerate(sorted(got)):
        res, rec = got[rid]
        assert rec["error"] == 0
        check_answer(oracle_mod, exported[i], res, 1e-4)
        solo = gpu.solve(gpu.Lp.generate(4, i), pdhg_tolerance=1e-4, **REL)
        assert solo.iterations == res.iterations, i
        np.testing.assert_array_equal(solo.col_value, res.col_value)
        np.testing.assert_array_equal(solo.row_dual, res.row_dual)


def test_admission_after_a_row_heavy_batch_does_not_stall(gpu):
    """One slot and a budget that holds either request alone but not the per-dimension maxima
    of both: after a row-heavy LP, an nnz-heavy one must still be admitted (the slot gives its
    buffers back) rather than wait forever."""
    base = gpu.Lp.preset(1, 3)
    row_heavy = gpu.Lp.generate_params(**dict(base, num_rows=50000, num_cols=1000,
                                              num_equality_rows=30000, degree_kind=0, degree_lo=5,
                                              fix_empty_rows=0))
    nnz_heavy = gpu.Lp.generate_params(**dict(base, num_rows=1000, num_cols=20000,
                                              num_equality_rows=600, degree_kind=0, degree_lo=20))

    def est(m, n, z):  # estimate_device_bytes of engine.hip, FP32 iterates, one instance
        return int(np.ceil(77.1 * z + 108.5 * m + 156.5 * n + 1664 + 262144))

    a, b = row_heavy.dims, nnz_heavy.dims
    alone = max(est(*a), est(*b))
    joint = est(max(a[0], b[0]), max(a[1], b[1]), max(a[2], b[2]))
    budget = (alone + joint) // 2
    assert alone < budget < joint
    srv = gpu.Server(slots=1, max_batch=1, memory_budget=budget, pdhg_tolerance=1e-4,
                     iteration_limit=200, **REL)
    dims = {srv.submit(row_heavy): a[:2]}
    got = _drain(srv, dims, timeout=60)
    dims2 = {srv.submit(nnz_heavy): b[:2]}
    got.update(_drain(srv, dims2, timeout=60))
    st = srv.stats()
    srv.close()
    assert st["completed"] == 2 and st["failed"] == 0 and st["releases"] >= 1
    assert st["peak_held"] <= budget
