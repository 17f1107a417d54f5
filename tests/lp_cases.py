"""Small LPs shared by the CPU and GPU tests.

The closed-form cases are the ones the reference's own PDHG tests solve
(tests/unit/test_pdhg.cpp:95-146), with the optimum stated in each entry. The
KKT-backwards generator builds an LP around a chosen primal-dual pair, so its optimum is
known exactly; it follows the construction described at tests/oracles/lp_generator.hpp:48-53
(min c'x s.t. Ax >= b, 0 <= x <= u), drawn with numpy's generator instead of std::mt19937_64.
"""
from __future__ import annotations

import numpy as np

INF = float("inf")

# (name, rows, row_lower, row_upper, cost, col_upper or None, sense, offset, optimum, x*)
CLOSED_FORM = [
    ("sum", [[1.0, 1.0]], [2.0], [INF], [1.0, 1.0], None, 1, 0.0, 2.0, None),
    ("equality", [[1.0, 1.0]], [4.0], [4.0], [1.0, 2.0], None, 1, 0.0, 4.0, [4.0, 0.0]),
    ("range", [[1.0]], [2.0], [6.0], [1.0], None, 1, 0.0, 2.0, [2.0]),
    ("column_upper", [[1.0]], [-INF], [100.0], [-1.0], [7.0], 1, 0.0, -7.0, [7.0]),
    ("maximise", [[1.0, 0.0], [0.0, 2.0], [3.0, 2.0]], [-INF, -INF, -INF], [4.0, 12.0, 18.0],
     [3.0, 5.0], None, -1, 0.0, 36.0, [2.0, 6.0]),
    ("offset", [[1.0]], [3.0], [INF], [1.0], None, 1, 10.0, 13.0, [3.0]),
]


def dense_to_lp(rows, row_lower, row_upper, cost, col_upper=None, sense=1, offset=0.0,
                col_lower=None) -> dict:
    a = np.asarray(rows, float).reshape(len(row_lower), len(cost))
    m, n = a.shape
    starts, idx, val = [0], [], []
    for j in range(n):
        for i in range(m):
            if a[i, j] != 0.0:
                idx.append(i)
                val.append(a[i, j])
        starts.append(len(idx))
    return {"num_rows": m, "num_cols": n, "col_starts": np.array(starts, np.int32),
            "row_indices": np.array(idx, np.int32), "values": np.array(val, float),
            "cost": np.asarray(cost, float),
            "col_lower": np.zeros(n) if col_lower is None else np.asarray(col_lower, float),
            "col_upper": np.full(n, INF) if col_upper is None else np.asarray(col_upper, float),
            "row_lower": np.asarray(row_lower, float), "row_upper": np.asarray(row_upper, float),
            "sense": sense, "offset": offset}


def closed_form(name: str) -> tuple[dict, float, list | None]:
    for case in CLOSED_FORM:
        if case[0] == name:
            _, rows, lo, hi, c, cu, sense, off, opt, xs = case
            return dense_to_lp(rows, lo, hi, c, cu, sense, off), opt, xs
    raise KeyError(name)


def kkt_lp(rng: np.random.Generator, m: int, n: int, density: float = 0.6,
           magnitude: int = 6) -> tuple[dict, float]:
    """An LP min c'x s.t. Ax >= b, 0 <= x <= u with a known optimum (integer data)."""
    a = rng.integers(-magnitude, magnitude + 1, size=(m, n)) * (rng.random((m, n)) < density)
    x = np.where(rng.random(n) < 0.6, rng.integers(0, 6, n), 0)
    active = rng.random(m) < 0.5
    y = np.where(active, rng.integers(0, 6, m), 0)
    d = np.where(x > 0, 0, rng.integers(0, 6, n))
    c = a.T @ y + d
    act = a @ x
    b = np.where(active, act, act - 1 - rng.integers(0, 6, m))
    upper = np.where(rng.random(n) < 0.35, x + rng.integers(0, 4, n), INF).astype(float)
    lp = dense_to_lp(a.astype(float), b.astype(float), np.full(m, INF), c.astype(float), upper)
    return lp, float(c @ x)


def csc_to_dense(lp: dict) -> np.ndarray:
    a = np.zeros((lp["num_rows"], lp["num_cols"]))
    cs, ri, v = lp["col_starts"], lp["row_indices"], lp["values"]
    for j in range(lp["num_cols"]):
        for p in range(cs[j], cs[j + 1]):
            a[ri[p], j] = v[p]
    return a


def kkt_backwards_from(lp: dict, seed: int, multiply) -> tuple[dict, float, np.ndarray, np.ndarray]:
    """The LP `lp` (its matrix and column bounds; rows with equal bounds stay equalities, the
    others become <= rows) with costs and right-hand sides chosen so that a drawn point is
    optimal (the construction of tests/oracles/lp_generator.hpp:48-53 on a given matrix):
    x* at its lower bound (30%), upper bound (20%) or inside; row multipliers free on
    equalities, <= 0 on active <= rows, 0 on inactive ones (slack U[0.5, 5]); reduced costs
    >= 0 at lower bounds, <= 0 at upper bounds, 0 inside; c = A^T y* + d*. Returns the LP, its
    optimum c'x*, x* and y* in the Solution's signs (row_dual). `multiply(lp, v, transpose)`
    is the FP64 oracle product."""
    rng = np.random.default_rng(seed)
    m, n = lp["num_rows"], lp["num_cols"]
    lo, up = np.asarray(lp["col_lower"], float), np.asarray(lp["col_upper"], float)
    assert np.all(np.isfinite(lo)) and np.all(np.isfinite(up)) and np.all(up > lo)
    r = rng.random(n)
    at_lo, at_up = r < 0.3, (r >= 0.3) & (r < 0.5)
    x = np.where(at_lo, lo, np.where(at_up, up, lo + (up - lo) * rng.uniform(0.2, 0.8, n)))
    act = multiply(lp, x, False)
    eq = np.asarray(lp["row_lower"]) == np.asarray(lp["row_upper"])
    active = ~eq & (rng.random(m) < 0.5)
    y = np.where(eq, rng.standard_normal(m), np.where(active, -np.abs(rng.standard_normal(m)), 0.0))
    row_lower = np.where(eq, act, -INF)
    row_upper = np.where(eq | active, act, act + rng.uniform(0.5, 5.0, m))
    d = np.where(at_lo, rng.uniform(0.1, 1.0, n), np.where(at_up, -rng.uniform(0.1, 1.0, n), 0.0))
    c = multiply(lp, y, True) + d
    out = dict(lp, cost=c, row_lower=row_lower, row_upper=row_upper, sense=1, offset=0.0)
    return out, float(c @ x), x, y
