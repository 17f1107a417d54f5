"""CPU pins of tests/spmv_order.py, the numpy restatement of the engine's per-row summation
order that the GPU probe tests hold the step kernels to."""
import numpy as np

from tests.spmv_order import (csr_rows, error_bound, fixed_order_sums, sliced_step_product,
                              step_product)


def random_csc(rng, m, n, per_col):
    cs, ri = [0], []
    for _ in range(n):
        k = int(per_col())
        ri.extend(np.sort(rng.choice(m, size=min(k, m), replace=False)))
        cs.append(len(ri))
    z = len(ri)
    return {"num_rows": m, "num_cols": n, "col_starts": np.array(cs, np.int32),
            "row_indices": np.array(ri, np.int32),
            "values": rng.standard_normal(z) * np.exp(rng.uniform(-3, 3, z))}


def dense(lp):
    a = np.zeros((lp["num_rows"], lp["num_cols"]))
    for j in range(lp["num_cols"]):
        for p in range(lp["col_starts"][j], lp["col_starts"][j + 1]):
            a[lp["row_indices"][p], j] = lp["values"][p]
    return a


def test_short_rows_are_the_oracle_serial_sums_bit_for_bit(oracle_mod):
    # columns of at most 64 entries: A^T y summed serially in CSC order is exactly the
    # oracle's gather (sparse.cpp:214-229, products rounded before each add)
    rng = np.random.default_rng(3)
    lp = random_csc(rng, 300, 200, lambda: rng.integers(0, 60))
    y = rng.standard_normal(300)
    got = step_product(lp, lp["values"], y, 2, np.float64)
    want = oracle_mod.multiply(dict(lp, cost=np.zeros(200), col_lower=np.zeros(200),
                                    col_upper=np.zeros(200), row_lower=np.zeros(300),
                                    row_upper=np.zeros(300)), y, transpose=True)
    np.testing.assert_array_equal(got, want)


def test_rows_in_lp_order_are_the_oracle_scatter_bit_for_bit(oracle_mod):
    # rows of A of at most 64 entries in the LP's column order: the oracle's column scatter
    # (sparse.cpp:202-212) adds each row's terms in the same order
    rng = np.random.default_rng(4)
    lp = random_csc(rng, 400, 300, lambda: rng.integers(0, 6))
    assert np.bincount(lp["row_indices"], minlength=400).max() <= 64
    x = rng.standard_normal(300)
    x[::7] = 0.0  # the scatter skips x_j == 0; adding +0 products changes nothing either
    got = step_product(lp, lp["values"], x, 1, np.float64)
    full = dict(lp, cost=np.zeros(300), col_lower=np.zeros(300), col_upper=np.zeros(300),
                row_lower=np.zeros(400), row_upper=np.zeros(400))
    np.testing.assert_array_equal(got, oracle_mod.multiply(full, x))


def test_long_rows_follow_the_lane_order():
    # one row of 300 terms: lane l sums terms l, l+64, ...; then the butterfly
    rng = np.random.default_rng(5)
    t = rng.standard_normal(300).astype(np.float32)
    lanes = np.zeros(64, np.float32)
    for k in range(300):
        lanes[k % 64] = np.float32(lanes[k % 64] + t[k])
    for o in (32, 16, 8, 4, 2, 1):
        lanes = lanes + lanes[np.arange(64) ^ o]
    got = fixed_order_sums(np.array([0, 300]), t)
    assert got.dtype == np.float32 and got[0] == lanes[0]
    assert abs(float(got[0]) - float(np.sum(t.astype(np.float64)))) <= 1e-4 * np.sum(np.abs(t))


def test_csr_rows_hold_lp_columns_ascending():
    rng = np.random.default_rng(6)
    lp = random_csc(rng, 50, 120, lambda: rng.integers(1, 8))
    ptr, cols, entry = csr_rows(lp)
    for i in range(50):
        assert np.all(np.diff(cols[ptr[i]:ptr[i + 1]]) > 0)
    assert np.all(lp["row_indices"][entry] == np.repeat(np.arange(50), np.diff(ptr)))
    x = rng.standard_normal(120)
    got = step_product(lp, lp["values"], x, 1, np.float64)
    np.testing.assert_allclose(got, dense(lp) @ x, rtol=1e-12, atol=1e-12)


def test_error_bound_holds_for_float32_in_any_order(oracle_mod):
    rng = np.random.default_rng(7)
    lp = random_csc(rng, 80, 2000, lambda: rng.integers(1, 5))
    full = dict(lp, cost=np.zeros(2000), col_lower=np.zeros(2000), col_upper=np.zeros(2000),
                row_lower=np.zeros(80), row_upper=np.zeros(80))
    x = rng.standard_normal(2000)
    got = step_product(lp, lp["values"], x, 1, np.float32).astype(np.float64)
    exact = dense(lp) @ x

    def mult(vals, vec, transpose):
        return oracle_mod.multiply(dict(full, values=vals), vec, transpose=transpose)

    bound = error_bound(lp, lp["values"], x, 1, 2.0 ** -24, mult)
    assert np.all(np.abs(got - exact) <= bound)
    # and the bound is not vacuous: float32 rounding is visible against it
    assert np.max(np.abs(got - exact) / bound) > 1e-3


def test_one_slice_is_the_unsliced_order_and_eight_are_slice_sums():
    # one slice is every row whole: the unsliced order bit for bit; eight slices are, per row,
    # the fixed-order sums of its entries in each column range, added in range order
    rng = np.random.default_rng(17)
    lp = random_csc(rng, 40, 400, lambda: rng.integers(1, 12))
    x = rng.standard_normal(400)
    np.testing.assert_array_equal(sliced_step_product(lp, lp["values"], x, 1, 1, np.float32),
                                  step_product(lp, lp["values"], x, 1, np.float32))
    y = rng.standard_normal(40)
    np.testing.assert_array_equal(sliced_step_product(lp, lp["values"], y, 2, 1, np.float32),
                                  step_product(lp, lp["values"], y, 2, np.float32))
    a = dense(lp).astype(np.float32)
    xf = x.astype(np.float32)
    got = sliced_step_product(lp, lp["values"], x, 1, 8, np.float32)
    bounds = [k * 400 // 8 for k in range(9)]
    for i in range(40):
        total = None
        for k in range(8):
            part = np.float32(0)
            for j in range(bounds[k], bounds[k + 1]):  # each slice holds < 64 entries of the row
                if a[i, j] != 0:
                    part = np.float32(part + np.float32(a[i, j] * xf[j]))
            total = part if total is None else np.float32(total + part)
        assert got[i] == total, i


def test_row_slices_are_slice_sums_of_each_column():
    # A^T y over 4 row slices: per column, the serial sums of its entries in each row range,
    # added in range order
    rng = np.random.default_rng(19)
    lp = random_csc(rng, 400, 30, lambda: rng.integers(1, 60))
    y = rng.standard_normal(400)
    got = sliced_step_product(lp, lp["values"], y, 2, 4, np.float32)
    a = dense(lp).astype(np.float32)
    yf = y.astype(np.float32)
    bounds = [k * 400 // 4 for k in range(5)]
    for j in range(30):
        total = None
        for k in range(4):
            part = np.float32(0)
            for i in range(bounds[k], bounds[k + 1]):  # at most 59 entries per column
                if a[i, j] != 0:
                    part = np.float32(part + np.float32(a[i, j] * yf[i]))
            total = part if total is None else np.float32(total + part)
        assert got[j] == total, j
