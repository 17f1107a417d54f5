"""MPS reader (fixed and free format) into the reference Model layout.

Semantics follow the reference's reader (src/io/mps_reader.cpp, src/io/mps_bounds.cpp):
rows N/E/L/G; the first N row is the objective and its RHS entry is the NEGATED objective
constant; RANGES widen rows as the MPS standard defines; BOUNDS UP/LO/FX/FR/MI/PL/BV;
columns default to [0, +inf); duplicate coefficients are summed and entries below 1e-11 in
magnitude are dropped (SparseMatrix::finalize, src/la/sparse.cpp:55-136); |bound| >= 1e30
is infinite (include/sankhya/types.hpp:32-45).
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

INF = math.inf


def _num(text: str) -> float:
    v = float(text)
    if math.isinf(v) or abs(v) >= 1e30:
        return INF if v > 0 else -INF
    return v


def read_mps(path: str | Path) -> dict:
    """Parse an MPS file into the dict layout of sih26_amd.Lp.export()."""
    row_kind: dict[str, str] = {}
    row_order: list[str] = []
    objective: str | None = None
    col_order: list[str] = []
    col_index: dict[str, int] = {}
    entries: dict[int, dict[str, float]] = {}
    cost: dict[int, float] = {}
    rhs: dict[str, float] = {}
    ranges: dict[str, float] = {}
    lower: dict[int, float] = {}
    upper: dict[int, float] = {}
    offset = 0.0
    sense = 1
    section = None
    for raw in Path(path).read_text().splitlines():
        if not raw.strip() or raw.startswith("*"):
            continue
        if not raw[0].isspace():
            head = raw.split()
            section = head[0].upper()
            if section == "OBJSENSE" and len(head) > 1:
                sense = -1 if head[1].upper().startswith("MAX") else 1
            continue
        f = raw.split()
        if section == "OBJSENSE":
            sense = -1 if f[0].upper().startswith("MAX") else 1
        elif section == "ROWS":
            kind, name = f[0].upper(), f[1]
            if kind == "N":
                if objective is None:
                    objective = name
                continue
            row_kind[name] = kind
            row_order.append(name)
        elif section == "COLUMNS":
            if "'MARKER'" in f:
                continue
            col = f[0]
            if col not in col_index:
                col_index[col] = len(col_order)
                col_order.append(col)
            j = col_index[col]
            for k in range(1, len(f) - 1, 2):
                row, val = f[k], float(f[k + 1])
                if row == objective:
                    cost[j] = cost.get(j, 0.0) + val
                elif row in row_kind:
                    e = entries.setdefault(j, {})
                    e[row] = e.get(row, 0.0) + val
        elif section == "RHS":
            items = f[1:] if len(f) % 2 == 1 else f
            for k in range(0, len(items) - 1, 2):
                row, val = items[k], float(items[k + 1])
                if row == objective:
                    offset = -val
                else:
                    rhs[row] = val
        elif section == "RANGES":
            items = f[1:] if len(f) % 2 == 1 else f
            for k in range(0, len(items) - 1, 2):
                ranges[items[k]] = float(items[k + 1])
        elif section == "BOUNDS":
            kind = f[0].upper()
            col = f[2] if len(f) >= 3 else f[1]
            val = _num(f[3]) if len(f) >= 4 else 0.0
            if col not in col_index:
                col_index[col] = len(col_order)
                col_order.append(col)
            j = col_index[col]
            if kind == "UP":
                upper[j] = val
            elif kind == "LO":
                lower[j] = val
            elif kind == "FX":
                lower[j] = upper[j] = val
            elif kind == "FR":
                lower[j], upper[j] = -INF, INF
            elif kind == "MI":
                lower[j] = -INF
            elif kind == "PL":
                upper[j] = INF
            elif kind == "BV":
                lower[j], upper[j] = 0.0, 1.0
            else:
                raise ValueError(f"unsupported bound type {kind}")
    m, n = len(row_order), len(col_order)
    ridx = {name: i for i, name in enumerate(row_order)}
    rl, ru = np.empty(m), np.empty(m)
    for i, name in enumerate(row_order):
        b = rhs.get(name, 0.0)
        kind = row_kind[name]
        lo, hi = {"E": (b, b), "L": (-INF, b), "G": (b, INF)}[kind]
        if name in ranges:
            r = ranges[name]
            if kind == "E":
                lo, hi = (b, b + abs(r)) if r >= 0 else (b - abs(r), b)
            elif kind == "L":
                lo = b - abs(r)
            else:
                hi = b + abs(r)
        rl[i], ru[i] = lo, hi
    starts, rows, vals = [0], [], []
    for j in range(n):
        for name, v in sorted(entries.get(j, {}).items(), key=lambda kv: ridx[kv[0]]):
            if not abs(v) < 1e-11:
                rows.append(ridx[name])
                vals.append(v)
        starts.append(len(rows))
    return {"num_rows": m, "num_cols": n, "col_starts": np.array(starts, np.int32),
            "row_indices": np.array(rows, np.int32), "values": np.array(vals, float),
            "cost": np.array([cost.get(j, 0.0) for j in range(n)]),
            "col_lower": np.array([lower.get(j, 0.0) for j in range(n)]),
            "col_upper": np.array([upper.get(j, INF) for j in range(n)]),
            "row_lower": rl, "row_upper": ru, "sense": sense, "offset": offset,
            "row_names": row_order, "col_names": col_order}
