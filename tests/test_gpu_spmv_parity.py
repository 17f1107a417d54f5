"""The step kernels' own sparse products and the device scaling, read back from the GPU and
held to the oracle (src/la/sparse.cpp:202-239, src/la/scaling.cpp:111-159).

Every row-block schedule is run: short-row tiles, long-row tiles (rows of more than 64
nonzeros) and vector tiles (one row above the tile capacity, forced with tile_nnz = 512), on
both tables (CSR(A) for A x-bar, CSR(A^T) for A^T y'), in FP32 and FP64, in the LP's column
order and in the most-referenced-first order. Three checks per product:
  - bit for bit the numpy restatement of the fixed per-row order (tests/spmv_order.py) applied
    to the scaled values the device holds;
  - within the a-priori rounding bound (k + 2) u sum |a_ij v_j| of the FP64 oracle product of
    the oracle's scaling (which is the reference's, tests/test_oracle.py);
  - the same bits at every tile capacity (a row's sum depends on the row alone).
"""
import numpy as np
import pytest

from tests.spmv_order import error_bound, sliced_step_product, step_product

pytestmark = pytest.mark.gpu

GOLDEN = "tests/golden/ref_scaling.npz"
GOLDEN_CASES = ["afiro", "c1_seed1", "c3_member30", "random_a", "random_b"]


def golden_lp(z, name):
    m, n, sense = (int(v) for v in z[f"{name}/in/dims"])
    d = {k: z[f"{name}/in/{k}"] for k in ("col_starts", "row_indices", "values", "cost",
                                             "col_lower", "col_upper", "row_lower", "row_upper")}
    d.update(num_rows=m, num_cols=n, sense=sense)
    return d


def to_lp(sk, d):
    return sk.Lp.from_arrays(d["num_rows"], d["num_cols"], d["col_starts"], d["row_indices"],
                             d["values"], d["cost"], d["col_lower"], d["col_upper"],
                             d["row_lower"], d["row_upper"], d.get("sense", 1), d.get("offset", 0.0))


def solved_batch(gpu, lp, **kw):
    """A one-LP batch after a short solve (the device set-up has run)."""
    b = gpu.Batch(iteration_limit=3, **kw)
    b.add(lp)
    b.upload()
    b.solve()
    return b


def power_law(gpu, seed=5):
    # power-law column degrees up to 2,000 at 2,000 x 5,000: CSR(A^T) rows of every kind
    return gpu.Lp.generate_params(**dict(gpu.Lp.preset(5, seed), num_rows=2000, num_cols=5000,
                                         num_equality_rows=1200, target_nnz=25000))


def long_rows(gpu, seed=9):
    # 20 rows x 4,000 columns, 5 nonzeros per column: every row of A has ~1,000 nonzeros
    return gpu.Lp.generate_params(**dict(gpu.Lp.preset(1, seed), num_rows=20, num_cols=4000,
                                         num_equality_rows=12))


@pytest.mark.parametrize("name", GOLDEN_CASES)
@pytest.mark.parametrize("precision,order", [("fp32", 2), ("fp64", 2), ("fp32", 1)])
def test_device_scaling_is_the_reference_bit_for_bit(gpu, name, precision, order):
    z = np.load(GOLDEN)
    d = golden_lp(z, name)
    b = solved_batch(gpu, to_lp(gpu, d), pdhg_precision=precision, column_order=order)
    s = b.probe_scaling(0)
    np.testing.assert_array_equal(s["row"], z[f"{name}/scaling/row"])
    np.testing.assert_array_equal(s["column"], z[f"{name}/scaling/column"])
    np.testing.assert_array_equal(s["values"], z[f"{name}/scaling/values"])


CASES = {
    # name: (lp factory, batch options)
    "c1": (lambda g: g.Lp.generate(1, 1), {}),
    "c3_member0": (lambda g: g.Lp.generate(3, 0), {}),
    "power_law_cap512": (power_law, {"tile_nnz": 512}),
    "power_law_hot_first": (power_law, {"tile_nnz": 512, "column_order": 1}),
    "long_rows_cap512": (long_rows, {"tile_nnz": 512}),
    "long_rows_default": (long_rows, {}),
    "power_law_fp64_cap512": (power_law, {"tile_nnz": 512, "pdhg_precision": "fp64"}),
    "long_rows_fp64_cap512": (long_rows, {"tile_nnz": 512, "pdhg_precision": "fp64"}),
}


@pytest.mark.parametrize("case", list(CASES))
def test_step_products_match_the_oracle(gpu, oracle_mod, case):
    make, opts = CASES[case]
    lp = make(gpu)
    d = lp.export()
    m, n = d["num_rows"], d["num_cols"]
    b = solved_batch(gpu, lp, **opts)
    fp64 = opts.get("pdhg_precision") == "fp64"
    dtype, unit = (np.float64, 2.0 ** -53) if fp64 else (np.float32, 2.0 ** -24)
    scaled = oracle_mod.build_scaling(d)["values"]
    np.testing.assert_array_equal(b.probe_scaling(0)["values"], scaled)
    order = b.column_order(0)
    assert sorted(order) == list(range(n))
    if opts.get("column_order") == 1:
        # the 2 MiB of columns with the most nonzeros first (largest first, ties in index
        # order), then the rest in the LP's order
        deg = np.diff(d["col_starts"])
        hot = min(n, (2 << 20) // (8 if fp64 else 4))
        head = sorted(range(n), key=lambda j: (-deg[j], j))[:hot]
        rest = sorted(set(range(n)) - set(head))
        np.testing.assert_array_equal(order, head + rest)
    rng = np.random.default_rng(11)
    x, y = rng.standard_normal(n), rng.standard_normal(m)

    def mult(vals, vec, transpose):
        return oracle_mod.multiply(dict(d, values=vals), vec, transpose=transpose)

    row_len = np.bincount(d["row_indices"], minlength=m)
    col_len = np.diff(d["col_starts"])
    for which, vec in ((1, x), (2, y)):
        got = b.probe_spmv(0, which, vec)
        want = step_product(d, scaled, vec, which, dtype).astype(np.float64)
        np.testing.assert_array_equal(got, want, err_msg=f"{case} product {which}: fixed order")
        exact = mult(scaled, vec, which == 2)
        bound = error_bound(d, scaled, vec, which, unit, mult)
        if fp64:
            bound = 2 * bound  # the oracle's own FP64 sum rounds too
        assert np.all(np.abs(got - exact) <= bound), (case, which, np.max(np.abs(got - exact) / bound))
    cap = opts.get("tile_nnz", 2560)
    if case.startswith("long_rows"):
        assert row_len.min() > 64  # the long-row path of A x-bar
        assert (row_len.max() > cap) == ("cap512" in case)  # vector tiles of A at 512
    if case.startswith("power_law"):
        # vector tiles (above the capacity) and long-row tiles of A^T
        assert col_len.max() > cap and ((col_len > 64) & (col_len <= cap)).any()


@pytest.mark.parametrize("make", [power_law, long_rows])
def test_tile_capacity_changes_no_row_sum(gpu, make):
    """A row's sum depends on the row alone: the same bits at every capacity, whether the row
    is a short-row, long-row or vector tile there."""
    lp = make(gpu)
    _, n, _ = lp.dims
    m = lp.dims[0]
    rng = np.random.default_rng(12)
    x, y = rng.standard_normal(n), rng.standard_normal(m)
    seen = None
    for cap in (512, 1024, 2560, 4096):
        b = solved_batch(gpu, lp, tile_nnz=cap)
        got = (b.probe_spmv(0, 1, x), b.probe_spmv(0, 2, y))
        if seen is None:
            seen = got
        else:
            np.testing.assert_array_equal(got[0], seen[0], err_msg=f"A x at capacity {cap}")
            np.testing.assert_array_equal(got[1], seen[1], err_msg=f"A^T y at capacity {cap}")


def test_batch_member_products_are_its_own(gpu, oracle_mod):
    """Probing one member of a batch gives its products alone (block-diagonal layout)."""
    lps = [gpu.Lp.generate(3, k) for k in (3, 4, 5)]
    b = gpu.Batch(iteration_limit=3, streams=1)
    for lp in lps:
        b.add(lp)
    b.upload()
    b.solve()
    for k, lp in enumerate(lps):
        d = lp.export()
        x = np.random.default_rng(k).standard_normal(d["num_cols"])
        scaled = oracle_mod.build_scaling(d)["values"]
        want = step_product(d, scaled, x, 1, np.float32).astype(np.float64)
        np.testing.assert_array_equal(b.probe_spmv(k, 1, x), want)


SLICED_CASES = {
    # sk_pdhg_options.column_slices = 1. A x-bar over 8 column slices: short-row tiles of the
    # stacked slices (power law), long-row tiles (rows of ~125 nonzeros per slice) and vector
    # tiles (the same rows above a capacity of 100). A^T y' over 4 row slices: the power law's
    # columns of up to 2,000 entries give long-row tiles, and vector tiles at 100
    "power_law": (power_law, {}),
    "power_law_cap512": (power_law, {"tile_nnz": 512}),
    "power_law_cap100": (power_law, {"tile_nnz": 100}),
    "long_rows": (long_rows, {}),
    "long_rows_cap100": (long_rows, {"tile_nnz": 100}),
    "power_law_fp64": (power_law, {"pdhg_precision": "fp64"}),
}


@pytest.mark.parametrize("case", list(SLICED_CASES))
def test_sliced_products_match_the_oracle(gpu, oracle_mod, case):
    """The sliced products, A x-bar over 8 column slices and A^T y' over 4 row slices: bit for
    bit the restatement of their order (each slice's row sums in the fixed per-row order, then
    the slices added in order), within the rounding bound of the FP64 oracle product; the
    device scaling is that of the unsliced engine."""
    make, opts = SLICED_CASES[case]
    lp = make(gpu)
    d = lp.export()
    m, n = d["num_rows"], d["num_cols"]
    b = solved_batch(gpu, lp, column_slices=1, **opts)
    assert b.stats()["column_slices"] == 8
    np.testing.assert_array_equal(b.column_order(0), np.arange(n))  # a sliced LP keeps its order
    fp64 = opts.get("pdhg_precision") == "fp64"
    dtype, unit = (np.float64, 2.0 ** -53) if fp64 else (np.float32, 2.0 ** -24)
    scaled = oracle_mod.build_scaling(d)["values"]
    np.testing.assert_array_equal(b.probe_scaling(0)["values"], scaled)
    rng = np.random.default_rng(13)
    x, y = rng.standard_normal(n), rng.standard_normal(m)

    def mult(vals, vec, transpose):
        return oracle_mod.multiply(dict(d, values=vals), vec, transpose=transpose)

    for which, vec, slices in ((1, x, 8), (2, y, 4)):
        got = b.probe_spmv(0, which, vec)
        want = sliced_step_product(d, scaled, vec, which, slices, dtype).astype(np.float64)
        np.testing.assert_array_equal(got, want, err_msg=f"{case} product {which}")
        exact = mult(scaled, vec, which == 2)
        bound = error_bound(d, scaled, vec, which, unit, mult) * (2 if fp64 else 1)
        assert np.all(np.abs(got - exact) <= bound), (which, np.max(np.abs(got - exact) / bound))
    if case == "power_law_cap100":
        cs, ri = np.asarray(d["col_starts"]), np.asarray(d["row_indices"])
        per = [np.diff(np.searchsorted(ri[cs[j]:cs[j + 1]], [k * m // 4 for k in range(5)]))
               for j in range(n)]
        assert max(v.max() for v in per) > 100  # a column's entries in one row slice: a vector tile
    if case.startswith("long_rows"):
        cs, ri = np.asarray(d["col_starts"]), np.asarray(d["row_indices"])
        lens = [np.bincount(ri[cs[k * n // 8]:cs[(k + 1) * n // 8]], minlength=m) for k in range(8)]
        assert min(v.min() for v in lens) > 64  # every row of every slice is a long row
        if "cap100" in case:
            assert max(v.max() for v in lens) > 100  # and some are vector tiles at 100


def test_slices_apply_to_a_batch_of_one_instance(gpu):
    """Slicing is per batch of one LP: a batch of two runs unsliced even when asked."""
    lps = [power_law(gpu, 5), power_law(gpu, 6)]
    b = gpu.Batch(iteration_limit=3, column_slices=1, streams=1)
    for lp in lps:
        b.add(lp)
    b.upload()
    b.solve()
    assert b.stats()["column_slices"] == 0
    one = solved_batch(gpu, lps[0], column_slices=2)
    assert one.stats()["column_slices"] == 0


def test_sliced_solve_meets_the_request_and_repeats(gpu, oracle_mod):
    """A solve with A x-bar over column slices: it meets the request as the oracle measures
    the answer, and a second solve gives the same bits (every sum has a fixed order)."""
    lp = power_law(gpu)
    d = lp.export()
    opts = dict(pdhg_tolerance=1e-4, pdhg_stop_at_request=True, pdhg_absolute_primal=False,
                column_slices=1)
    a, b = gpu.solve(lp, **opts), gpu.solve(lp, **opts)
    assert a.converged, a.info["message"]
    ev = oracle_mod.evaluate(d, a.col_value, -d["sense"] * a.row_dual)
    assert ev["worst"] <= 1e-4 * (1 + 1e-9), ev
    assert a.iterations == b.iterations
    np.testing.assert_array_equal(a.col_value, b.col_value)
    np.testing.assert_array_equal(a.row_dual, b.row_dual)
