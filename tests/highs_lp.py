"""An independent LP solver for the presolve tests: HiGHS through scipy.optimize.linprog (a
dual-simplex vertex, the kind of answer the reference's presolve tests postsolve). Test
infrastructure only. Answers come back in the reference Solution's signs: col_dual = c - A^T
row_dual, in the model's own objective sense."""
from __future__ import annotations

import numpy as np
from scipy.optimize import linprog
from scipy.sparse import csc_matrix

STATUS = {0: "optimal", 2: "infeasible", 3: "unbounded"}


def solve(lp: dict) -> dict:
    m, n = lp["num_rows"], lp["num_cols"]
    sense = lp.get("sense", 1)
    c = sense * np.asarray(lp["cost"], float)
    lo, hi = np.asarray(lp["col_lower"], float), np.asarray(lp["col_upper"], float)
    bounds = list(zip(np.where(np.isfinite(lo), lo, None), np.where(np.isfinite(hi), hi, None)))
    if n == 0:  # nothing to choose: the rows hold iff zero lies in their bounds
        ok = bool(np.all((lp["row_lower"] <= 1e-9) & (lp["row_upper"] >= -1e-9)))
        return {"status": "optimal" if ok else "infeasible", "x": np.zeros(0), "row_dual": np.zeros(m),
                "col_dual": np.zeros(0), "objective": lp.get("offset", 0.0)}
    a = csc_matrix((lp["values"], lp["row_indices"], lp["col_starts"]), shape=(m, n)).tocsr()
    rl, ru = np.asarray(lp["row_lower"], float), np.asarray(lp["row_upper"], float)
    eq = np.isfinite(rl) & np.isfinite(ru) & (rl == ru)
    up = ~eq & np.isfinite(ru)
    low = ~eq & np.isfinite(rl)
    kw = {}
    if eq.any():
        kw.update(A_eq=a[eq], b_eq=rl[eq])
    ub_rows = [a[up], -a[low]]
    if up.any() or low.any():
        from scipy.sparse import vstack
        kw.update(A_ub=vstack(ub_rows).tocsr(), b_ub=np.concatenate([ru[up], -rl[low]]))
    res = linprog(c, bounds=bounds, method="highs-ds", **kw)
    out = {"status": STATUS.get(res.status, f"highs status {res.status}"), "message": res.message}
    if res.status in (2, 3):
        # HiGHS may report "infeasible or unbounded" as either; the feasibility problem (zero
        # cost) separates the two
        feas = linprog(np.zeros(n), bounds=bounds, method="highs-ds", **kw)
        out["status"] = "unbounded" if feas.status == 0 else "infeasible"
    if res.status != 0:
        return out
    y = np.zeros(m)
    if eq.any():
        y[eq] = res.eqlin.marginals
    if up.any() or low.any():
        mu = res.ineqlin.marginals
        nu = int(up.sum())
        y[up] += mu[:nu]
        y[low] -= mu[nu:]
    x = np.asarray(res.x, float)
    y = sense * y  # back to the model's own sense
    d = np.asarray(lp["cost"], float) - a.T @ y
    out.update(x=x, row_dual=y, col_dual=d,
               objective=float(lp.get("offset", 0.0) + np.dot(lp["cost"], x)))
    return out
