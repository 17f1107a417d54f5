"""Independent LP answer checker: a restatement of the reference's checker
(tools/verify_solution.py) for linear programs. TEST INFRASTRUCTURE ONLY: it judges answers
the engine wrote; nothing in the product imports it.

The reference's checker trusts nothing the solver says: it recomputes activities, the
objective, reduced costs and the duality gap from (model, x, y) and holds each to the claim
the file's status makes. This module does the same from the .sol text (written by
sk_write_solution in the reference's format, src/io/writer.cpp:119-227) and the model arrays
(the layout of sih26_amd.Lp.export() / tests/mps_ref.read_mps), so it can run where the
reference cannot travel. Check names, thresholds and the order of checks follow
verify_solution.py (line numbers cited per check); the model is taken as arrays rather than
re-read from MPS. Integer columns, quadratic objectives and infeasibility certificates are
out of scope (the engine answers LPs and emits no certificates).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

INF = math.inf
PRIMAL_TOL = 1e-7   # verify_solution.py:41-44 (tolerances.hpp defaults)
DUAL_TOL = 1e-7
DUALITY_TOL = 1e-9

# verify_solution.py:702-715
WITH_A_POINT = ("optimal", "feasible", "unbounded", "iteration_limit", "time_limit", "node_limit")
ASSERTS_FEASIBILITY = ("optimal", "feasible", "unbounded")


@dataclass
class SolFile:
    header: dict = field(default_factory=dict)
    columns: dict = field(default_factory=dict)  # name -> (value, reduced cost)
    rows: dict = field(default_factory=dict)     # name -> (activity, dual)

    @property
    def status(self) -> str:
        return self.header.get("status", "unknown")

    def number(self, key: str):
        try:
            return float(self.header[key])
        except (KeyError, ValueError):
            return None


def _fields(line: str) -> list[str]:
    """Whitespace fields, with a leading double-quoted name holding backslash escapes (the
    writer quotes names that contain whitespace, writer.cpp:65-82)."""
    if not line.startswith('"'):
        return line.split()
    name, i = [], 1
    while i < len(line):
        ch = line[i]
        if ch == "\\" and i + 1 < len(line):
            name.append(line[i + 1])
            i += 2
        elif ch == '"':
            return ["".join(name)] + line[i + 1:].split()
        else:
            name.append(ch)
            i += 1
    return line.split()


def read_sol(path: str | Path) -> SolFile:
    sol, block = SolFile(), ""
    for raw in Path(path).read_text(errors="replace").splitlines():
        line = raw.strip()
        if not line or line.startswith("#"):
            continue
        f = _fields(line)
        if f[0] in ("begin", "end"):
            block = f[1] if f[0] == "begin" and len(f) > 1 else ""
        elif block == "columns" and len(f) >= 3:
            sol.columns[f[0]] = (float(f[1]), float(f[2]))
        elif block == "rows" and len(f) >= 3:
            sol.rows[f[0]] = (float(f[1]), float(f[2]))
        elif not block and len(f) >= 2:
            sol.header[f[0]] = " ".join(f[1:])
    return sol


@dataclass
class Report:
    lines: list = field(default_factory=list)  # (ok, name, detail)

    def check(self, ok: bool, name: str, detail: str = "") -> None:
        self.lines.append((bool(ok), name, detail))

    def note(self, name: str, detail: str) -> None:
        self.lines.append((True, name, detail))

    @property
    def failures(self) -> int:
        return sum(1 for ok, _, _ in self.lines if not ok)

    @property
    def verified(self) -> bool:
        return self.failures == 0

    def outcome(self) -> list[tuple[str, bool]]:
        return [(name, ok) for ok, name, _ in self.lines]


def _bound(v: float) -> float:
    return INF if v >= 1e30 else (-INF if v <= -1e30 else v)


def verify(model: dict, sol: SolFile, primal_tol: float = PRIMAL_TOL, dual_tol: float = DUAL_TOL,
           duality_tol: float = DUALITY_TOL) -> Report:
    """The checks of verify_solution.py:718-1147 for an LP, in the same order."""
    rep = Report()
    m, n = int(model["num_rows"]), int(model["num_cols"])
    sigma = -1.0 if int(model.get("sense", 1)) < 0 else 1.0
    col_names = list(model.get("col_names") or [f"C{j}" for j in range(n)])
    row_names = list(model.get("row_names") or [f"R{i}" for i in range(m)])
    starts = np.asarray(model["col_starts"], np.int64)
    ridx = np.asarray(model["row_indices"], np.int64)
    vals = np.asarray(model["values"], float)
    cost = np.asarray(model["cost"], float)
    cl = np.array([_bound(v) for v in model["col_lower"]])
    cu = np.array([_bound(v) for v in model["col_upper"]])
    rl = np.array([_bound(v) for v in model["row_lower"]])
    ru = np.array([_bound(v) for v in model["row_upper"]])
    offset = float(model.get("offset", 0.0))
    col_of = np.repeat(np.arange(n), np.diff(starts))

    # a verdict that claims no point is not checked as if it did (:745-754)
    if sol.status not in WITH_A_POINT and sol.status != "infeasible_or_unbounded":
        rep.note("verdict", f"status is {sol.status}, which claims no point")
        return rep

    # structure (:763-774)
    missing_c = [c for c in col_names if c not in sol.columns]
    missing_r = [r for r in row_names if r not in sol.rows]
    rep.check(not missing_c and not missing_r, "structure",
              f"missing {len(missing_c)} columns, {len(missing_r)} rows")
    if missing_c or missing_r:
        return rep
    x = np.array([sol.columns[c][0] for c in col_names])
    asserts = sol.status in ASSERTS_FEASIBILITY

    def primal_check(worst_rel: float, worst_abs: float, name: str) -> None:
        # :779-794: held to feasibility when asserted, else to the file's own stated figure
        if asserts:
            rep.check(worst_rel <= primal_tol, name, f"worst {worst_abs:.3e} ({worst_rel:.3e})")
            return
        stated = sol.number("primal_infeasibility")
        if stated is None:
            rep.note(name, "no feasibility asserted and none stated")
            return
        rep.check(worst_abs <= stated * (1.0 + 1e-6) + primal_tol, name + " matches the stated",
                  f"worst {worst_abs:.3e}, stated {stated:.3e}")

    # column bounds, scaled by the value's own magnitude (:796-805)
    with np.errstate(invalid="ignore"):
        viol = np.maximum(np.maximum(cl - x, x - cu), 0.0)
    rel = viol / np.maximum(1.0, np.abs(x))
    k = int(np.argmax(rel)) if n else 0
    primal_check(float(rel[k]) if n else 0.0, float(viol[k]) if n else 0.0, "column bounds")
    rep.note("integrality", "no integer columns")  # :808-822

    # row activity recomputed, judged against the largest term of each row (:824-866)
    terms = vals * x[col_of]
    activity = np.zeros(m)
    np.add.at(activity, ridx, np.where(x[col_of] == 0.0, 0.0, terms))
    scale = np.ones(m)
    np.maximum.at(scale, ridx, np.where(x[col_of] == 0.0, 0.0, np.abs(terms)))
    with np.errstate(invalid="ignore"):
        rviol = np.maximum(np.maximum(rl - activity, activity - ru), 0.0)
    rrel = rviol / scale
    k = int(np.argmax(rrel)) if m else 0
    primal_check(float(rrel[k]) if m else 0.0, float(rviol[k]) if m else 0.0, "row activity")
    stated_act = np.array([sol.rows[r][0] for r in row_names])
    gap_act = float(np.max(np.abs(activity - stated_act))) if m else 0.0
    rep.check(gap_act <= 1e-6, "activity agreement", f"max |ours - solver's| = {gap_act:.3e}")

    # objective recomputed (:877-892)
    objective = offset + float(cost @ x)
    claimed = sol.number("objective")
    if claimed is None:
        rep.check(False, "objective", "no objective stated")
    else:
        rep.check(abs(objective - claimed) <= 1e-9 * max(1.0, abs(objective)), "objective",
                  f"recomputed {objective:.12e}, stated {claimed:.12e}")
    if sol.status != "optimal":  # duality is a test of optimality only (:894-901)
        rep.note("duality", f"skipped: status is {sol.status}")
        return rep

    # minimise space (:950-951)
    y = sigma * np.array([sol.rows[r][1] for r in row_names])
    d = sigma * np.array([sol.columns[c][1] for c in col_names])
    c_min = sigma * cost
    yterm = vals * y[ridx]
    aty = np.zeros(n)
    np.add.at(aty, col_of, yterm)
    col_scale = np.maximum(1.0, np.abs(c_min))
    np.maximum.at(col_scale, col_of, np.abs(yterm))

    # reduced costs d = c - A^T y, relative to the terms (:972-987)
    diff = np.abs(c_min - aty - d)
    k = int(np.argmax(diff / col_scale)) if n else 0
    rep.check((float((diff / col_scale)[k]) if n else 0.0) <= 1e-6, "reduced costs",
              f"max |c - A^T y - d| = {float(diff[k]) if n else 0.0:.3e}")

    def sign_violation(mult, lower, upper):
        # a multiplier pushing against a bound that does not exist (:989-1014)
        out = np.zeros_like(mult)
        free_side = lower != upper
        out = np.where(free_side & (mult > 0) & ~np.isfinite(lower), mult, out)
        out = np.where(free_side & (mult < 0) & ~np.isfinite(upper), -mult, out)
        return out

    cv = sign_violation(d, cl, cu) / col_scale
    rep.check((float(np.max(cv)) if n else 0.0) <= dual_tol, "dual feasibility (columns)",
              f"worst {float(np.max(cv)) if n else 0.0:.3e}")
    dual_norm = max(1.0, float(np.max(np.abs(y))) if m else 0.0)
    rv = sign_violation(y, rl, ru) / dual_norm
    rep.check((float(np.max(rv)) if m else 0.0) <= dual_tol, "dual feasibility (rows)",
              f"worst {float(np.max(rv)) if m else 0.0:.3e}")

    def compl(mult, value, lower, upper):
        # |multiplier| * nearest slack; |multiplier| alone where no side is finite (:1038-1064)
        with np.errstate(invalid="ignore"):
            sl = np.where(np.isfinite(lower), value - lower, INF)
            su = np.where(np.isfinite(upper), upper - value, INF)
        slack = np.minimum(sl, su)
        prod = np.where(np.isfinite(slack), np.abs(mult) * np.where(np.isfinite(slack), slack, 0.0),
                        np.abs(mult))
        return np.where(lower == upper, 0.0, prod)

    worst_c = 0.0
    if m:
        worst_c = max(worst_c, float(np.max(compl(y, activity, rl, ru))))
    if n:
        worst_c = max(worst_c, float(np.max(compl(d, x, cl, cu))))
    rep.check(worst_c <= 1e-6, "complementary slackness", f"worst {worst_c:.3e}")

    def contribution(mult, lower, upper):
        # the bound a multiplier's sign prices (pdhg.cpp:175-199 / verify_solution.py)
        return np.where(mult > 0, np.where(np.isfinite(lower), mult * np.where(np.isfinite(lower), lower, 0.0), 0.0),
                        np.where(mult < 0, np.where(np.isfinite(upper), mult * np.where(np.isfinite(upper), upper, 0.0), 0.0), 0.0))

    dual_min = float(np.sum(contribution(y, rl, ru))) + float(np.sum(contribution(d, cl, cu)))

    def share(mult, value, lower, upper, mscale):
        # each item's share of the gap that the per-item checks accepted (:1090-1133)
        priced = np.where(mult > 0, lower, upper)
        with np.errstate(invalid="ignore"):
            term = np.abs(mult) * np.where(np.isfinite(priced), np.abs(value - np.where(np.isfinite(priced), priced, 0.0)), np.abs(value))
            nl = np.where(np.isfinite(lower), np.abs(value - np.where(np.isfinite(lower), lower, 0.0)), INF)
            nu = np.where(np.isfinite(upper), np.abs(value - np.where(np.isfinite(upper), upper, 0.0)), INF)
        nearest = np.minimum(nl, nu)
        small = np.abs(mult) <= dual_tol * mscale
        limited = np.where(np.isfinite(nearest), np.minimum(term, np.abs(mult) * np.where(np.isfinite(nearest), nearest, 0.0)), 0.0)
        return np.where(mult == 0.0, 0.0, np.where(small, term, limited))

    accounted = float(np.sum(share(d, x, cl, cu, col_scale))) + \
        float(np.sum(share(y, activity, rl, ru, np.full(m, dual_norm))))
    dual_objective = sigma * dual_min + offset
    gap = abs(objective - dual_objective)
    rep.check(gap <= duality_tol * max(1.0, abs(objective)) + accounted, "strong duality",
              f"primal {objective:.12e} dual {dual_objective:.12e} gap {gap:.3e}, accounted {accounted:.3e}")
    return rep
