"""The engine's fixed per-row summation order, restated in numpy for the probe tests.

A row of at most 64 terms (kLongRow) is summed serially from zero in its stored order; a
longer row by 64 lanes, lane l adding terms l, l + 64, l + 128, ... in order from zero, then a
butterfly over the lanes (partner lane l ^ o for o = 32, 16, ..., 1), whose lane-0 value is the
sum. Every addition and product is one IEEE operation in `dtype`, so the result is the device's
bit for bit when the device rounds each product before adding it. Stored order: CSR(A^T) rows
(columns of A) hold rows of A ascending; CSR(A) rows hold the LP's columns ascending (also where
the device keeps the columns in another order).
"""
from __future__ import annotations

import numpy as np

LONG_ROW = 64


def fixed_order_sums(ptr: np.ndarray, terms: np.ndarray) -> np.ndarray:
    dtype = terms.dtype
    ptr = np.asarray(ptr, np.int64)
    rows = ptr.size - 1
    lens = np.diff(ptr)
    out = np.zeros(rows, dtype)
    short = lens <= LONG_ROW
    for k in range(LONG_ROW):
        sel = np.nonzero(short & (lens > k))[0]
        if sel.size == 0:
            break
        out[sel] = out[sel] + terms[ptr[sel] + k]
    long_rows = np.nonzero(~short)[0]
    if long_rows.size:
        acc = np.zeros((long_rows.size, LONG_ROW), dtype)
        start, length = ptr[long_rows], lens[long_rows]
        lane = np.arange(LONG_ROW)
        for r0 in range(0, int(length.max()), LONG_ROW):
            pos = r0 + lane[None, :]
            ok = pos < length[:, None]
            vals = np.where(ok, terms[np.where(ok, start[:, None] + pos, 0)], dtype.type(0))
            # lanes past the row's end keep their sum: x + 0 is x exactly
            acc = np.where(ok, acc + vals, acc)
        for o in (32, 16, 8, 4, 2, 1):
            acc = acc + acc[:, lane ^ o]
        out[long_rows] = acc[:, 0]
    return out


def csr_rows(lp: dict):
    """CSR(A) as the device holds it: each row's entries in the LP's column order, whatever
    order the device keeps the columns in. Returns the row pointers, the column of each entry
    and the index of each entry in the LP's CSC arrays."""
    m, n = lp["num_rows"], lp["num_cols"]
    cs = np.asarray(lp["col_starts"], np.int64)
    ri = np.asarray(lp["row_indices"], np.int64)
    cols = np.repeat(np.arange(n), np.diff(cs))
    order = np.lexsort((cols, ri))
    ptr = np.zeros(m + 1, np.int64)
    np.cumsum(np.bincount(ri, minlength=m), out=ptr[1:])
    return ptr, cols[order], order


def step_product(lp: dict, scaled_values: np.ndarray, vec: np.ndarray, which: int,
                 dtype) -> np.ndarray:
    """What the step kernels compute: which 1 A-hat vec (rows), which 2 A-hat^T vec (columns),
    with the scaled values and vec rounded to `dtype` first, in LP order."""
    vals = np.asarray(scaled_values, np.float64).astype(dtype)
    v = np.asarray(vec, np.float64).astype(dtype)
    if which == 2:
        ri = np.asarray(lp["row_indices"], np.int64)
        return fixed_order_sums(lp["col_starts"], vals * v[ri])
    ptr, cols, entry = csr_rows(lp)
    return fixed_order_sums(ptr, vals[entry] * v[cols])


def sliced_step_product(lp: dict, scaled_values: np.ndarray, vec: np.ndarray, which: int,
                        slices: int, dtype) -> np.ndarray:
    """What the step kernels compute over slices (sk_pdhg_options.column_slices): which 1,
    A-hat vec over column slices (slice k: columns [k n / slices, (k + 1) n / slices) of the
    LP's order, integer division); which 2, A-hat^T vec over row slices (rows [k m / slices,
    (k + 1) m / slices)). Each row's entries in a slice are summed in the fixed per-row order
    (a row of the stacked slices), and a row's slice sums are added in slice order, all in
    `dtype`."""
    vals = np.asarray(scaled_values, np.float64).astype(dtype)
    v = np.asarray(vec, np.float64).astype(dtype)
    m, n = lp["num_rows"], lp["num_cols"]
    if which == 1:
        ptr, gather, entry = csr_rows(lp)
        terms = vals[entry] * v[gather]
        count, rows = n, m
    else:
        ptr = np.asarray(lp["col_starts"], np.int64)
        gather = np.asarray(lp["row_indices"], np.int64)
        terms = vals * v[gather]
        count, rows = m, n
    owner = np.repeat(np.arange(rows), np.diff(ptr))
    bounds = (np.arange(slices + 1, dtype=np.int64) * count) // slices
    slice_of = np.searchsorted(bounds, gather, side="right") - 1
    out = None
    for k in range(slices):
        sel = slice_of == k  # stored order: each row's entries of the slice stay one run, in order
        sptr = np.zeros(rows + 1, np.int64)
        np.cumsum(np.bincount(owner[sel], minlength=rows), out=sptr[1:])
        part = fixed_order_sums(sptr, terms[sel])
        out = part if out is None else out + part
    return out


def error_bound(lp: dict, scaled_values: np.ndarray, vec: np.ndarray, which: int, unit: float,
                multiply) -> np.ndarray:
    """A-priori bound on |computed - exact| for products rounded from rounded inputs and summed
    in any order: (k + 2) u sum |a_ij v_j| per row of k terms (Higham, ch. 3), with 1% slack
    for the second-order terms. `multiply(values, vec, transpose)` is the FP64 oracle product."""
    mag = multiply(np.abs(np.asarray(scaled_values, np.float64)), np.abs(vec), which == 2)
    if which == 2:
        k = np.diff(np.asarray(lp["col_starts"], np.int64))
    else:
        k = np.bincount(np.asarray(lp["row_indices"], np.int64), minlength=lp["num_rows"])
    return 1.01 * (k + 2) * unit * mag + 1e-300
