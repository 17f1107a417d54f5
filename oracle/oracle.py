"""ctypes wrapper of the FP64 CPU oracle (liboracle.so) and of the reference harness
(_ref/libsankhya_ref.so, built only where the reference sources exist).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker or the timed CPU baseline, never by the product.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path
from typing import Any

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF = HERE / "_ref" / "libsankhya_ref.so"
REFERENCE_ROOT = Path("/root/reference")

i32, i64, f64, vp = C.c_int32, C.c_int64, C.c_double, C.c_void_p


class Model(C.Structure):
    _fields_ = [("num_rows", i32), ("num_cols", i32), ("sense", i32), ("objective_offset", f64),
                ("col_starts", vp), ("row_indices", vp), ("values", vp), ("col_cost", vp),
                ("col_lower", vp), ("col_upper", vp), ("row_lower", vp), ("row_upper", vp)]


class Residuals(C.Structure):
    _fields_ = [(k, f64) for k in ("primal", "dual", "gap", "primal_objective", "dual_objective",
                                   "absolute_primal", "absolute_dual", "gap_as_verified",
                                   "complementarity")]


class Options(C.Structure):
    _fields_ = [("tolerance", f64), ("iteration_limit", i64), ("time_limit", f64),
                ("restart", i32), ("stop_at_request", i32), ("random_seed", i64),
                ("absolute_primal", i32)]


class Solution(C.Structure):
    _fields_ = [("status", i32), ("converged", i32), ("iterations", i64), ("restarts", i64),
                ("evaluations", i64), ("objective", f64), ("dual_bound", f64),
                ("solve_seconds", f64), ("setup_seconds", f64), ("spectral_norm", f64),
                ("final_residuals", Residuals), ("primal_infeasibility", f64),
                ("primal_infeasibility_scaled", f64), ("dual_infeasibility", f64),
                ("dual_infeasibility_scaled", f64), ("complementarity_violation", f64),
                ("absolute_gap", f64), ("relative_gap", f64), ("col_value", vp),
                ("col_dual", vp), ("row_dual", vp), ("row_activity", vp)]


def build(ref: bool | None = None, arch_flags: str = "") -> None:
    """Compile liboracle.so (and the reference harness when the sources are present)."""
    cmd = ["make", "-s", "-C", str(HERE), "-B", "liboracle.so"]
    if arch_flags:
        cmd.append(f"ARCH_FLAGS={arch_flags}")
    subprocess.run(cmd, check=True)
    if ref is None:
        ref = (REFERENCE_ROOT / "src" / "la" / "scaling.cpp").exists()
    if ref:
        subprocess.run(["make", "-s", "-C", str(HERE), "-B", "ref"], check=True)


_lib = None
_ref = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build(ref=False)
        _lib = C.CDLL(str(LIB))
        _lib.ok_evaluate.restype = Residuals
        _lib.ok_cpu_dot.restype = f64
        _lib.ok_spectral_norm.restype = f64
        _lib.ok_project.restype = f64
        _lib.ok_project.argtypes = [f64, f64, f64]
        _lib.ok_is_finite_bound.argtypes = [f64]
        _lib.ok_evaluate.argtypes = [C.POINTER(Model), vp, f64, f64, vp, vp, vp, vp]
        _lib.ok_solve_pdhg.argtypes = [C.POINTER(Model), C.POINTER(Options), C.POINTER(Solution)]
        _lib.ok_recompute_quality.argtypes = [C.POINTER(Model), C.POINTER(Solution)]
        _lib.ok_spectral_norm.argtypes = [i32, i32, vp, vp, vp, i32, C.c_uint32]
        _lib.ok_mt19937_uniform.argtypes = [C.c_uint32, i32, vp]
        _lib.ok_mt19937_raw.argtypes = [C.c_uint32, i32, vp]
        _lib.ok_cpu_axpy.argtypes = [i32, f64, vp, vp]
        _lib.ok_cpu_dot.argtypes = [i32, vp, vp]
        _lib.ok_cpu_primal_step.argtypes = [i32, f64, vp, vp, vp, vp, vp, vp, vp]
        _lib.ok_cpu_spmv_csc.argtypes = [i32, i32, vp, vp, vp, vp, vp, f64, i32]
        _lib.ok_compare_vectors.argtypes = [i32, vp, vp, f64, C.POINTER(f64), C.POINTER(f64),
                                            C.POINTER(i32)]
        _lib.ok_multiply.argtypes = [i32, i32, vp, vp, vp, vp, vp]
        _lib.ok_transpose_multiply.argtypes = [i32, i32, vp, vp, vp, vp, vp]
        _lib.ok_set_threads.argtypes = [i32]
        _lib.ok_build_scaling.argtypes = [C.POINTER(Model), vp, i32] + [vp] * 10
    return _lib


def ref() -> C.CDLL | None:
    """The reference's own scaling/sparse code, or None where it was not built."""
    global _ref
    if _ref is None and REF.exists():
        _ref = C.CDLL(str(REF))
        _ref.ref_spectral_norm.restype = f64
        _ref.ref_spectral_norm.argtypes = [i32, i32, vp, vp, vp, i32, C.c_uint32]
        _ref.ref_build_scaling.restype = i32
        _ref.ref_build_scaling.argtypes = [i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, i32] + [vp] * 10
        _ref.ref_multiply.argtypes = [i32, i32, vp, vp, vp, vp, vp]
        _ref.ref_transpose_multiply.argtypes = [i32, i32, vp, vp, vp, vp, vp]
    return _ref


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def arrays(lp: dict[str, Any]) -> dict[str, Any]:
    """Contiguous typed copies of an LP dict (keys as sih26_amd.Lp.export())."""
    out = dict(lp)
    for k in ("col_starts", "row_indices"):
        out[k] = np.ascontiguousarray(lp[k], np.int32)
    for k in ("values", "cost", "col_lower", "col_upper", "row_lower", "row_upper"):
        out[k] = np.ascontiguousarray(lp[k], np.float64)
    out.setdefault("sense", 1)
    out.setdefault("offset", 0.0)
    return out


def model(lp: dict[str, Any]) -> tuple[Model, dict[str, Any]]:
    a = arrays(lp)
    mdl = Model(a["num_rows"], a["num_cols"], a["sense"], a["offset"], _p(a["col_starts"]),
                _p(a["row_indices"]), _p(a["values"]), _p(a["cost"]), _p(a["col_lower"]),
                _p(a["col_upper"]), _p(a["row_lower"]), _p(a["row_upper"]))
    return mdl, a  # keep `a` alive while `mdl` is used


def solve_pdhg(lp: dict[str, Any], tolerance: float = 1e-4, iteration_limit: int = -1,
               time_limit: float = 1.7976931348623157e308, restart: bool = True,
               stop_at_request: bool = False, random_seed: int = 0,
               absolute_primal: bool = True) -> dict[str, Any]:
    mdl, a = model(lp)
    m, n = a["num_rows"], a["num_cols"]
    x, d, yd, act = np.zeros(n), np.zeros(n), np.zeros(m), np.zeros(m)
    sol = Solution()
    sol.col_value, sol.col_dual, sol.row_dual, sol.row_activity = _p(x), _p(d), _p(yd), _p(act)
    opts = Options(tolerance, iteration_limit, time_limit, int(restart), int(stop_at_request),
                   random_seed, int(absolute_primal))
    lib().ok_solve_pdhg(C.byref(mdl), C.byref(opts), C.byref(sol))
    out = {f: getattr(sol, f) for f, _ in Solution._fields_ if f not in
           ("col_value", "col_dual", "row_dual", "row_activity", "final_residuals")}
    out["final"] = {f: getattr(sol.final_residuals, f) for f, _ in Residuals._fields_}
    out.update(col_value=x, col_dual=d, row_dual=yd, row_activity=act)
    return out


QUALITY_FIELDS = ("objective", "primal_infeasibility", "primal_infeasibility_scaled",
                  "dual_infeasibility", "dual_infeasibility_scaled", "complementarity_violation",
                  "absolute_gap", "relative_gap")


def recompute_quality(lp: dict[str, Any], col_value: np.ndarray, row_dual: np.ndarray,
                      col_dual: np.ndarray, dual_bound: float) -> dict[str, Any]:
    """model.cpp:228-422 (LP case) on a given answer in the Solution's signs: the quality
    measures and the recomputed row activities."""
    mdl, a = model(lp)
    x = np.ascontiguousarray(col_value, np.float64)
    yd = np.ascontiguousarray(row_dual, np.float64)
    d = np.ascontiguousarray(col_dual, np.float64)
    act = np.zeros(a["num_rows"])
    sol = Solution()
    sol.col_value, sol.col_dual, sol.row_dual, sol.row_activity = _p(x), _p(d), _p(yd), _p(act)
    sol.dual_bound = dual_bound
    lib().ok_recompute_quality(C.byref(mdl), C.byref(sol))
    out = {f: getattr(sol, f) for f in QUALITY_FIELDS}
    out["row_activity"] = act
    return out


def evaluate(lp: dict[str, Any], x: np.ndarray, y: np.ndarray) -> dict[str, Any]:
    """pdhg.cpp:143-263 on the unscaled model; y uses the engine's sign (row_dual = -sense y)."""
    mdl, a = model(lp)
    m, n = a["num_rows"], a["num_cols"]
    cost = a["sense"] * a["cost"]
    cost_norm = float(np.sqrt(np.sum(cost * cost)))
    b = np.where(np.isfinite(a["row_lower"]) & (np.abs(a["row_lower"]) < 1e30), a["row_lower"],
                 np.where(np.isfinite(a["row_upper"]) & (np.abs(a["row_upper"]) < 1e30),
                          a["row_upper"], 0.0))
    bound_norm = 0.0
    for v in b:  # the reference's sequential order
        bound_norm += v * v
    bound_norm = float(np.sqrt(bound_norm))
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    act, red = np.zeros(m), np.zeros(n)
    r = lib().ok_evaluate(C.byref(mdl), _p(np.ascontiguousarray(cost)), bound_norm, cost_norm,
                          _p(x), _p(y), _p(act), _p(red))
    out = {f: getattr(r, f) for f, _ in Residuals._fields_}
    out["worst"] = max(out["primal"], out["dual"], out["gap"])
    out["activity"] = act
    out["reduced"] = red
    return out


def build_scaling(lp: dict[str, Any], ruiz: int = 10, use_ref: bool = False) -> dict[str, Any]:
    mdl, a = model(lp)
    m, n, z = a["num_rows"], a["num_cols"], int(a["col_starts"][-1])
    cost = np.ascontiguousarray(a["sense"] * a["cost"])
    out = {k: np.zeros(s) for k, s in (("row", m), ("column", n), ("values", z), ("cost", n),
                                       ("col_lower", n), ("col_upper", n), ("row_lower", m),
                                       ("row_upper", m))}
    mn, mx = C.c_double(), C.c_double()
    ptrs = [_p(out[k]) for k in ("row", "column", "values", "cost", "col_lower", "col_upper",
                                 "row_lower", "row_upper")]
    if use_ref:
        r = ref()
        if r is None:
            raise RuntimeError("reference harness not built")
        kept = r.ref_build_scaling(m, n, _p(a["col_starts"]), _p(a["row_indices"]), _p(a["values"]),
                                   _p(a["col_lower"]), _p(a["col_upper"]), _p(a["row_lower"]),
                                   _p(a["row_upper"]), _p(cost), ruiz, *ptrs, C.byref(mn), C.byref(mx))
        out["kept"] = kept
    else:
        lib().ok_build_scaling(C.byref(mdl), _p(cost), ruiz, *ptrs, C.byref(mn), C.byref(mx))
    out["min_abs"], out["max_abs"] = mn.value, mx.value
    return out


def spectral_norm(lp: dict[str, Any], values: np.ndarray | None = None, iterations: int = 30,
                  seed: int = 1, use_ref: bool = False) -> float:
    a = arrays(lp)
    v = a["values"] if values is None else np.ascontiguousarray(values, np.float64)
    fn = ref().ref_spectral_norm if use_ref else lib().ok_spectral_norm
    return fn(a["num_rows"], a["num_cols"], _p(a["col_starts"]), _p(a["row_indices"]), _p(v),
              iterations, seed)


def set_threads(threads: int) -> None:
    """Threads of the A^T x gather (sparse.cpp:216-220); results do not depend on it."""
    lib().ok_set_threads(int(threads))


def multiply(lp: dict[str, Any], x: np.ndarray, transpose: bool = False, use_ref: bool = False) -> np.ndarray:
    a = arrays(lp)
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(a["num_cols"] if transpose else a["num_rows"])
    L = ref() if use_ref else lib()
    fn = (L.ref_transpose_multiply if transpose else L.ref_multiply) if use_ref else \
        (L.ok_transpose_multiply if transpose else L.ok_multiply)
    fn(a["num_rows"], a["num_cols"], _p(a["col_starts"]), _p(a["row_indices"]), _p(a["values"]),
       _p(x), _p(y))
    return y
