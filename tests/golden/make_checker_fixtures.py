"""Writes the answer-checker fixtures under tests/golden/checker/: models (MPS), answers (.sol,
written by this library's sk_write_solution) and the reference checker's verdict on each.

    python tests/golden/make_checker_fixtures.py      # here only: needs /root/reference

The answers are the FP64 oracle's solutions of a few LPs, plus deliberately wrong variants
(a shifted column value, a flipped row dual, a wrong objective, a limit status that
understates its infeasibility). The verdicts come from running the reference's independent
checker (tools/verify_solution.py; stdlib only) on each model/answer pair in this container;
they are data: the checks' names and pass/fail outcomes. tests/test_checker.py holds
oracle/checker.py to the same outcomes. Nothing of the reference is copied or kept.
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import json
import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
OUT = HERE / "checker"
REFERENCE_CHECKER = Path("/root/reference/tools/verify_solution.py")
sys.path.insert(0, str(ROOT))

import sih26_amd as sk  # noqa: E402
from sih26_amd import _ffi  # noqa: E402
from tests.mps_ref import read_mps  # noqa: E402
from oracle import oracle  # noqa: E402
from tests.lp_cases import closed_form  # noqa: E402


def fmt(v: float) -> str:
    return repr(float(v))


def write_mps(path: Path, d: dict, name: str) -> None:
    """Free-format MPS of an LP dict (rows R<i>, columns C<j>, objective row COST)."""
    m, n = d["num_rows"], d["num_cols"]
    inf = math.inf
    lines = [f"NAME {name}"]
    if d.get("sense", 1) < 0:
        lines += ["OBJSENSE", "    MAX"]
    lines.append("ROWS")
    lines.append(" N COST")
    kinds, rhs, rng = [], {}, {}
    for i in range(m):
        lo, hi = d["row_lower"][i], d["row_upper"][i]
        if lo == hi:
            kinds.append("E"); rhs[i] = lo
        elif lo == -inf:
            kinds.append("L"); rhs[i] = hi
        elif hi == inf:
            kinds.append("G"); rhs[i] = lo
        else:
            kinds.append("G"); rhs[i] = lo; rng[i] = hi - lo
        lines.append(f" {kinds[-1]} R{i}")
    lines.append("COLUMNS")
    for j in range(n):
        if d["cost"][j] != 0.0:
            lines.append(f"    C{j} COST {fmt(d['cost'][j])}")
        for p in range(d["col_starts"][j], d["col_starts"][j + 1]):
            lines.append(f"    C{j} R{d['row_indices'][p]} {fmt(d['values'][p])}")
    lines.append("RHS")
    if d.get("offset", 0.0) != 0.0:
        lines.append(f"    RHS COST {fmt(-d['offset'])}")
    for i, v in rhs.items():
        if v != 0.0:
            lines.append(f"    RHS R{i} {fmt(v)}")
    if rng:
        lines.append("RANGES")
        for i, v in rng.items():
            lines.append(f"    RNG R{i} {fmt(v)}")
    lines.append("BOUNDS")
    for j in range(n):
        lo, hi = d["col_lower"][j], d["col_upper"][j]
        if lo == -inf and hi == inf:
            lines.append(f" FR BND C{j}")
            continue
        if lo != 0.0:
            lines.append(f" MI BND C{j}" if lo == -inf else f" LO BND C{j} {fmt(lo)}")
        if hi != inf:
            lines.append(f" UP BND C{j} {fmt(hi)}")
    lines.append("ENDATA")
    path.write_text("\n".join(lines) + "\n")


def lp_handle(d: dict) -> sk.Lp:
    return sk.Lp.from_arrays(d["num_rows"], d["num_cols"], d["col_starts"], d["row_indices"],
                             d["values"], d["cost"], d["col_lower"], d["col_upper"],
                             d["row_lower"], d["row_upper"], d.get("sense", 1), d.get("offset", 0.0))


def write_answer(path: Path, d: dict, r: dict, x, yd, dd, act, status: int, objective: float,
                 primal_inf: float) -> None:
    res = _ffi.PdhgResult()
    res.status = status
    res.converged = 1 if status in (1, 5) else 0
    res.iterations = int(r["iterations"])
    res.objective = objective
    res.dual_bound = r["dual_bound"]
    res.primal_infeasibility = primal_inf
    res.dual_infeasibility = r["dual_infeasibility"]
    res.algorithm = b"pdhg-cpu"
    p = lambda a: C.c_void_p(np.ascontiguousarray(a, np.float64).ctypes.data)  # noqa: E731
    arrays = [np.ascontiguousarray(a, np.float64) for a in (x, yd, dd, act)]
    names_c = d.get("col_names")
    names_r = d.get("row_names")
    cn = (C.c_char_p * len(names_c))(*[s.encode() for s in names_c]) if names_c else None
    rn = (C.c_char_p * len(names_r))(*[s.encode() for s in names_r]) if names_r else None
    lp = lp_handle(d)
    st = sk.lib().sk_write_solution(str(path).encode(), lp._h, C.byref(res), *[p(a) for a in arrays],
                                    C.cast(cn, C.c_void_p) if cn else None,
                                    C.cast(rn, C.c_void_p) if rn else None, d.get("name", "").encode())
    assert st == 0, sk.lib().sk_gpu_last_error()


def main() -> int:
    spec = importlib.util.spec_from_file_location("reference_checker", REFERENCE_CHECKER)
    ref = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ref)
    oracle.build(ref=False)
    OUT.mkdir(exist_ok=True)
    models = {}
    for name in ("afiro", "sc50a", "blend"):
        mps = HERE / "netlib" / f"{name}.mps"
        d = read_mps(mps)
        d["name"] = name
        models[name] = (d, mps)
    for name in ("equality", "range", "sum"):
        d, _, _ = closed_form(name)
        d["name"] = f"closed_{name}"
        mps = OUT / f"closed_{name}.mps"
        write_mps(mps, d, d["name"])
        models[d["name"]] = (read_mps(mps) | {"name": d["name"]}, mps)
    g = sk.Lp.generate_params(seed=7, num_rows=30, num_cols=60, num_equality_rows=18,
                              degree_kind=0, degree_lo=3, degree_hi=3, value_kind=0,
                              value_lo=0.1, value_hi=2.0, negative_probability=0.2,
                              upper_lo=1.0, upper_hi=100.0, slack_lo=0.0, slack_hi=5.0,
                              cost_lo=-1.0, cost_hi=1.0, fix_empty_rows=1).export()
    mps = OUT / "blending_30x60.mps"
    write_mps(mps, g, "blending_30x60")
    models["blending_30x60"] = (read_mps(mps) | {"name": "blending_30x60"}, mps)

    verdicts = {}
    for name, (d, mps) in models.items():
        r = oracle.solve_pdhg(d, tolerance=1e-10, iteration_limit=400000)
        x, yd, dd, act = r["col_value"], r["row_dual"], r["col_dual"], r["row_activity"]
        variants = {"good": (x, yd, dd, act, r["status"], r["objective"], r["primal_infeasibility"])}
        j = int(np.argmax(np.abs(x)))
        xs = x.copy()
        xs[j] += 1e-3 * max(1.0, abs(x[j]))
        variants["shifted_column"] = (xs, yd, dd, act, r["status"], r["objective"], r["primal_infeasibility"])
        ineq = [i for i in range(d["num_rows"]) if d["row_lower"][i] != d["row_upper"][i]]
        if ineq:
            i = max(ineq, key=lambda k: abs(yd[k]))
            yf = yd.copy()
            yf[i] = -yf[i] if yf[i] != 0.0 else 1.0
            variants["flipped_row_dual"] = (x, yf, dd, act, r["status"], r["objective"], r["primal_infeasibility"])
        variants["wrong_objective"] = (x, yd, dd, act, r["status"],
                                       r["objective"] + 1e-6 * max(1.0, abs(r["objective"])),
                                       r["primal_infeasibility"])
        variants["limit_understated"] = (xs, yd, dd, act, 6, r["objective"], 0.0)
        for tag, (vx, vy, vd, va, status, obj, pinf) in variants.items():
            sol = OUT / f"{name}.{tag}.sol"
            write_answer(sol, d, r, vx, vy, vd, va, status, obj, pinf)
            report = ref.verify(ref.parse_mps(mps), ref.parse_sol(sol), ref.DEFAULT_PRIMAL_TOL,
                                ref.DEFAULT_DUAL_TOL, ref.DEFAULT_INTEGER_TOL,
                                ref.DEFAULT_DUALITY_TOL)
            verdicts[f"{name}.{tag}"] = {
                "model": str(mps.relative_to(HERE)), "sol": str(sol.relative_to(HERE)),
                "checks": [[check, bool(ok)] for ok, check, _ in report.lines],
                "verified": report.failures == 0}
            print(f"{name}.{tag}: {'VERIFIED' if report.failures == 0 else 'REJECTED'}")
    (OUT / "verdicts.json").write_text(json.dumps(verdicts, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
