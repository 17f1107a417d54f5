"""Python view of the C ABI (include/sk_gpu.h): LP handles, options, single solves, batches.

Every call goes through libsk_gpu.so; there is no other execution path. Arrays cross as
numpy buffers (int32 indices, float64 values).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from . import _ffi

_lib: C.CDLL | None = None

STATUS_NAMES = {0: "not_solved", 1: "optimal", 2: "infeasible", 3: "unbounded",
                4: "infeasible_or_unbounded", 5: "feasible", 6: "iteration_limit",
                7: "time_limit", 8: "node_limit", 9: "numerical_error", 10: "model_error"}


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = _ffi.load()
    return _lib


class SkError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        text = lib().sk_gpu_status_string(status).decode()
        detail = lib().sk_gpu_last_error().decode()
        super().__init__(f"{where}: {text} ({status}): {detail}")


def check(status: int, where: str) -> None:
    if status != 0:
        raise SkError(status, where)


def _ptr(a: np.ndarray | None) -> C.c_void_p | None:
    return None if a is None else C.c_void_p(a.ctypes.data)


def gpu_available() -> bool:
    return bool(lib().sk_gpu_is_available())


def device_info(device: int = 0) -> dict[str, Any]:
    info = _ffi.DeviceInfo()
    check(lib().sk_gpu_get_device_info(device, C.byref(info)), "sk_gpu_get_device_info")
    return {f: (getattr(info, f).decode() if isinstance(getattr(info, f), bytes) else getattr(info, f))
            for f, _ in info._fields_}


class Lp:
    """An LP held by the library (reference Model layout: CSC, two-sided rows)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @classmethod
    def from_arrays(cls, num_rows: int, num_cols: int, col_starts, row_indices, values, cost,
                    col_lower, col_upper, row_lower, row_upper, sense: int = 1,
                    offset: float = 0.0) -> "Lp":
        arrs = [np.ascontiguousarray(col_starts, np.int32), np.ascontiguousarray(row_indices, np.int32),
                np.ascontiguousarray(values, np.float64), np.ascontiguousarray(cost, np.float64),
                np.ascontiguousarray(col_lower, np.float64), np.ascontiguousarray(col_upper, np.float64),
                np.ascontiguousarray(row_lower, np.float64), np.ascontiguousarray(row_upper, np.float64)]
        h = C.c_void_p()
        check(lib().sk_lp_create(num_rows, num_cols, sense, offset, *[_ptr(a) for a in arrs],
                                 C.byref(h)), "sk_lp_create")
        return cls(h)

    @classmethod
    def from_dense(cls, rows, row_lower, row_upper, cost, col_lower=None, col_upper=None,
                   sense: int = 1, offset: float = 0.0) -> "Lp":
        a = np.asarray(rows, np.float64).reshape(len(row_lower), len(cost))
        m, n = a.shape
        starts, idx, val = [0], [], []
        for j in range(n):
            for i in range(m):
                if a[i, j] != 0.0:
                    idx.append(i)
                    val.append(a[i, j])
            starts.append(len(idx))
        lo = np.zeros(n) if col_lower is None else np.asarray(col_lower, float)
        hi = np.full(n, np.inf) if col_upper is None else np.asarray(col_upper, float)
        return cls.from_arrays(m, n, starts, idx, val, cost, lo, hi, row_lower, row_upper, sense, offset)

    @classmethod
    def read_mps(cls, path) -> "Lp":
        """An MPS file read by the library (sk_lp_read_mps); `integer_columns` and `names()`
        describe what the file said beyond the LP."""
        h = C.c_void_p()
        k = C.c_int32()
        check(lib().sk_lp_read_mps(str(path).encode(), C.byref(h), C.byref(k)), "sk_lp_read_mps")
        lp = cls(h)
        lp.integer_columns = k.value
        return lp

    def write_mps(self, path) -> None:
        """Free-format MPS that read_mps gives back as the same arrays (sk_lp_write_mps)."""
        check(lib().sk_lp_write_mps(self._h, str(path).encode()), "sk_lp_write_mps")

    def save(self, path) -> None:
        """The binary instance cache (sk_lp_save)."""
        check(lib().sk_lp_save(self._h, str(path).encode()), "sk_lp_save")

    @classmethod
    def load(cls, path) -> "Lp":
        """An LP from the binary instance cache (sk_lp_load)."""
        h = C.c_void_p()
        check(lib().sk_lp_load(str(path).encode(), C.byref(h)), "sk_lp_load")
        return cls(h)

    def names(self) -> tuple[str, list[str], list[str]]:
        """(model name, row names, column names) of an LP read from a file."""
        m, n, _ = self.dims
        get = lambda kind, i: (lib().sk_lp_name(self._h, kind, i) or b"").decode()  # noqa: E731
        return get(2, 0), [get(0, i) for i in range(m)], [get(1, j) for j in range(n)]

    @classmethod
    def generate(cls, config: int, instance: int) -> "Lp":
        h = C.c_void_p()
        check(lib().sk_lp_generate(config, instance, C.byref(h)), "sk_lp_generate")
        return cls(h)

    @staticmethod
    def preset(config: int, instance: int) -> dict[str, Any]:
        p = _ffi.GenParams()
        check(lib().sk_gen_preset(config, instance, C.byref(p)), "sk_gen_preset")
        return {f: getattr(p, f) for f, _ in p._fields_}

    @classmethod
    def generate_params(cls, **params) -> "Lp":
        base = cls.preset(1, 1)
        base.update(params)
        p = _ffi.GenParams(**base)
        h = C.c_void_p()
        check(lib().sk_lp_generate_params(C.byref(p), C.byref(h)), "sk_lp_generate_params")
        return cls(h)

    @property
    def dims(self) -> tuple[int, int, int]:
        m, n, z = C.c_int32(), C.c_int32(), C.c_int64()
        check(lib().sk_lp_dims(self._h, C.byref(m), C.byref(n), C.byref(z)), "sk_lp_dims")
        return m.value, n.value, z.value

    def export(self) -> dict[str, Any]:
        m, n, z = self.dims
        out = {"num_rows": m, "num_cols": n,
               "col_starts": np.empty(n + 1, np.int32), "row_indices": np.empty(z, np.int32),
               "values": np.empty(z), "cost": np.empty(n), "col_lower": np.empty(n),
               "col_upper": np.empty(n), "row_lower": np.empty(m), "row_upper": np.empty(m)}
        sense, off = C.c_int32(), C.c_double()
        check(lib().sk_lp_export(self._h, C.byref(sense), C.byref(off),
                                 *[_ptr(out[k]) for k in ("col_starts", "row_indices", "values", "cost",
                                                          "col_lower", "col_upper", "row_lower",
                                                          "row_upper")]), "sk_lp_export")
        out["sense"] = sense.value
        out["offset"] = off.value
        return out

    def close(self) -> None:
        if self._h:
            lib().sk_lp_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_options(**kw) -> _ffi.PdhgOptions:
    """Options with the reference defaults; keyword names are the struct fields
    (tolerance, iteration_limit, ...) or reference option names (pdhg_tolerance, ...)."""
    o = _ffi.PdhgOptions()
    lib().sk_pdhg_default_options(C.byref(o))
    fields = {f for f, _ in o._fields_}
    for k, v in kw.items():
        if k in fields and not isinstance(v, str):
            setattr(o, k, v)
        else:
            val = ("true" if v else "false") if isinstance(v, bool) else str(v)
            check(lib().sk_pdhg_set_option(C.byref(o), k.encode(), val.encode()), f"option {k}")
    return o


def adaptive_config(**kw) -> _ffi.AdaptiveConfig:
    """AdaptiveConfig (sk_adaptive_config) with the reference defaults; keywords are struct
    fields or the registry's adaptive_* option names."""
    c = _ffi.AdaptiveConfig()
    lib().sk_adaptive_default_config(C.byref(c))
    fields = {f for f, _ in c._fields_}
    for k, v in kw.items():
        if k in fields:
            setattr(c, k, v)
        else:
            val = ("true" if v else "false") if isinstance(v, bool) else str(v)
            check(lib().sk_adaptive_set_option(C.byref(c), k.encode(), val.encode()), f"option {k}")
    return c


def analyze_problem(lp: "Lp", integer_columns: int = 0,
                    quadratic: bool = False) -> _ffi.ProblemCharacteristics:
    """analyze_problem (src/core/adaptive.cpp:32-81) of an LP, with the HIP device's facts."""
    ch = _ffi.ProblemCharacteristics()
    check(lib().sk_analyze_problem(lp._h, integer_columns, int(quadratic), C.byref(ch)),
          "sk_analyze_problem")
    return ch


BACKENDS = {0: "auto", 1: "cpu", 2: "gpu", 3: "hybrid"}


def select_backend(chars: _ffi.ProblemCharacteristics, options: _ffi.PdhgOptions | None = None,
                   config: _ffi.AdaptiveConfig | None = None) -> dict[str, Any]:
    """select_backend (src/core/adaptive.cpp:98-232): which engine should run."""
    d = _ffi.BackendDecision()
    check(lib().sk_select_backend(C.byref(chars), C.byref(options) if options else None,
                                  C.byref(config) if config else None, C.byref(d)),
          "sk_select_backend")
    return {"selected": BACKENDS[d.selected_backend], "requested": BACKENDS[d.requested_mode],
            "valid": bool(d.is_valid), "reason": d.primary_reason.decode(),
            "explanation": d.explanation.decode(), "details": d.detail_count}


@dataclass
class SolveResult:
    status: int
    status_name: str
    info: dict[str, Any]
    col_value: np.ndarray
    row_dual: np.ndarray
    col_dual: np.ndarray
    row_activity: np.ndarray
    extra: dict[str, Any] = field(default_factory=dict)

    def __getattr__(self, name):
        info = self.__dict__.get("info", {})
        if name in info:
            return info[name]
        raise AttributeError(name)


def _result(res: _ffi.PdhgResult, x, yd, d, act) -> SolveResult:
    info = {f: getattr(res, f) for f, _ in res._fields_}
    info["algorithm"] = info["algorithm"].decode()
    info["message"] = info["message"].decode()
    return SolveResult(res.status, STATUS_NAMES.get(res.status, "?"), info, x, yd, d, act,
                       extra={"raw": res})


def write_solution(path, lp: Lp, result: SolveResult, model_name: str | None = None,
                   col_names=None, row_names=None) -> None:
    """The reference's .sol file for one answer (sk_write_solution)."""
    def names(seq):
        if seq is None:
            return None
        arr = (C.c_char_p * len(seq))(*[s.encode() for s in seq])
        return C.cast(arr, C.c_void_p), arr
    cn, rn = names(col_names), names(row_names)
    check(lib().sk_write_solution(str(path).encode(), lp._h, C.byref(result.extra["raw"]),
                                  _ptr(result.col_value), _ptr(result.row_dual),
                                  _ptr(result.col_dual), _ptr(result.row_activity),
                                  cn[0] if cn else None, rn[0] if rn else None,
                                  (model_name or "").encode()), "sk_write_solution")


def solve(lp: Lp, options: _ffi.PdhgOptions | None = None, **kw) -> SolveResult:
    """The drop-in of pdhg::solve_pdhg on one LP (runs on the GPU)."""
    o = options if options is not None else make_options(**kw)
    m, n, _ = lp.dims
    x, d, yd, act = np.zeros(n), np.zeros(n), np.zeros(m), np.zeros(m)
    res = _ffi.PdhgResult()
    check(lib().sk_pdhg_solve(lp._h, C.byref(o), C.byref(res), _ptr(x), _ptr(yd), _ptr(d), _ptr(act)),
          "sk_pdhg_solve")
    return _result(res, x, yd, d, act)


def ipm_options(**kw) -> _ffi.IpmOptions:
    """Interior-point options with the reference defaults; keywords are the struct fields
    (iteration_limit, start_margin, max_seconds, max_factor_nonzeros, scaling)."""
    o = _ffi.IpmOptions()
    lib().sk_ipm_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def solve_ipm(lp: Lp, warm: SolveResult | None = None, **kw) -> SolveResult:
    """ipm::solve_ipm on one LP, warm-started from `warm`'s point when given."""
    o = ipm_options(**kw)
    m, n, _ = lp.dims
    x, d, yd, act = np.zeros(n), np.zeros(n), np.zeros(m), np.zeros(m)
    res = _ffi.PdhgResult()
    w = (_ptr(warm.col_value), _ptr(warm.row_dual), _ptr(warm.col_dual)) if warm else (None,) * 3
    check(lib().sk_ipm_solve(lp._h, C.byref(o), *w, C.byref(res), _ptr(x), _ptr(yd), _ptr(d),
                             _ptr(act)), "sk_ipm_solve")
    return _result(res, x, yd, d, act)


def polish(lp: Lp, first: SolveResult, **kw) -> SolveResult:
    """The reference's PDHG-then-interior-point finish of one answer (sk_pdhg_polish)."""
    o = ipm_options(**kw)
    m, n, _ = lp.dims
    x, d, yd, act = np.zeros(n), np.zeros(n), np.zeros(m), np.zeros(m)
    res = _ffi.PdhgResult()
    fx = np.ascontiguousarray(first.col_value, np.float64)
    fy = np.ascontiguousarray(first.row_dual, np.float64)
    fd = np.ascontiguousarray(first.col_dual, np.float64)
    check(lib().sk_pdhg_polish(lp._h, C.byref(o), C.byref(first.extra["raw"]), _ptr(fx), _ptr(fy),
                               _ptr(fd), C.byref(res), _ptr(x), _ptr(yd), _ptr(d), _ptr(act)),
          "sk_pdhg_polish")
    return _result(res, x, yd, d, act)


def refine(lp: Lp, first: SolveResult, **kw) -> SolveResult:
    """sk_pdhg_refine: the FP64 engine warm-started from `first` (the finish of an LP above the
    interior point's factor cap)."""
    m, n, _ = lp.dims
    o = make_options(**kw)
    x, d, yd, act = np.zeros(n), np.zeros(n), np.zeros(m), np.zeros(m)
    res = _ffi.PdhgResult()
    check(lib().sk_pdhg_refine(lp._h, C.byref(o), C.byref(first.extra["raw"]),
                               _ptr(np.ascontiguousarray(first.col_value, np.float64)),
                               _ptr(np.ascontiguousarray(first.row_dual, np.float64)),
                               C.byref(res), _ptr(x), _ptr(yd), _ptr(d), _ptr(act)),
          "sk_pdhg_refine")
    return _result(res, x, yd, d, act)


PRESOLVE_KINDS = ("empty_row", "redundant_row", "fixed_column", "empty_column", "singleton_row",
                  "forcing_row", "free_column_singleton", "doubleton_equation")


class Presolved:
    """presolve::presolve on one LP (sk_presolve, host code): the reduced LP, the records, and
    postsolve of a reduced answer back to the original LP."""

    def __init__(self, lp: Lp, integer_columns=None, feasibility: float = 1e-7):
        self._h = C.c_void_p()
        mask = None if integer_columns is None else np.ascontiguousarray(integer_columns, np.uint8)
        check(lib().sk_presolve(lp._h, _ptr(mask), feasibility, C.byref(self._h)), "sk_presolve")
        self.original = lp
        info = _ffi.PresolveInfo()
        check(lib().sk_presolve_get_info(self._h, C.byref(info)), "sk_presolve_get_info")
        self.info = {f: getattr(info, f) for f, _ in info._fields_ if f != "by_kind"}
        self.info["by_kind"] = {k: info.by_kind[i] for i, k in enumerate(PRESOLVE_KINDS)}
        self.message = lib().sk_presolve_message(self._h).decode()
        self.proved_infeasible = bool(info.proved_infeasible)

    @property
    def reduced(self) -> Lp:
        h = C.c_void_p()
        check(lib().sk_presolve_reduced_lp(self._h, C.byref(h)), "sk_presolve_reduced_lp")
        return Lp(h)

    def maps(self) -> tuple[np.ndarray, np.ndarray]:
        cols = np.empty(self.info["reduced_cols"], np.int32)
        rows = np.empty(self.info["reduced_rows"], np.int32)
        check(lib().sk_presolve_maps(self._h, _ptr(cols), _ptr(rows)), "sk_presolve_maps")
        return cols, rows

    def records(self) -> list[dict[str, Any]]:
        out = []
        for k in range(self.info["records"]):
            r = _ffi.PresolveRecord()
            check(lib().sk_presolve_get_record(self._h, k, C.byref(r)), "sk_presolve_get_record")
            d = {f: getattr(r, f) for f, _ in r._fields_}
            d["kind"] = PRESOLVE_KINDS[r.kind]
            out.append(d)
        return out

    def postsolve(self, col_value, row_dual, col_dual) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """A reduced answer (Solution signs) back to the original LP: (x, row_dual, col_dual)."""
        m, n, _ = self.original.dims
        x, y, d = np.zeros(n), np.zeros(m), np.zeros(n)
        arrs = [np.ascontiguousarray(a, np.float64) for a in (col_value, row_dual, col_dual)]
        check(lib().sk_postsolve(self._h, self.original._h, *[_ptr(a) for a in arrs], _ptr(x),
                                 _ptr(y), _ptr(d)), "sk_postsolve")
        return x, y, d

    def close(self) -> None:
        if self._h:
            lib().sk_presolve_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def presolve(lp: Lp, integer_columns=None, feasibility: float = 1e-7) -> Presolved:
    return Presolved(lp, integer_columns, feasibility)


class Batch:
    """Many LPs solved together in one block-diagonal device solve."""

    def __init__(self, options: _ffi.PdhgOptions | None = None, **kw):
        self.options = options if options is not None else make_options(**kw)
        h = C.c_void_p()
        check(lib().sk_batch_create(C.byref(self.options), C.byref(h)), "sk_batch_create")
        self._h = h
        self._lps: list[Lp] = []

    def add(self, lp: Lp) -> int:
        k = C.c_int32()
        check(lib().sk_batch_add(self._h, lp._h, C.byref(k)), "sk_batch_add")
        self._lps.append(lp)
        return k.value

    def warm_start(self, k: int, col_value, row_dual, primal_weight: float = 0.0,
                   step_size: float = 0.0) -> None:
        """Instance k starts from (col_value, row_dual) (Solution signs; NaN keeps an entry's
        cold start); call before upload()."""
        m, n, _ = self._lps[k].dims
        x = np.ascontiguousarray(col_value, np.float64)
        y = np.ascontiguousarray(row_dual, np.float64)
        if x.shape != (n,) or y.shape != (m,):
            raise ValueError("warm_start: lengths do not match the LP")
        check(lib().sk_batch_warm_start(self._h, k, _ptr(x), _ptr(y), primal_weight, step_size),
              "sk_batch_warm_start")

    def upload(self) -> None:
        check(lib().sk_batch_upload(self._h), "sk_batch_upload")

    def solve(self) -> None:
        check(lib().sk_batch_solve(self._h), "sk_batch_solve")

    def result(self, k: int) -> SolveResult:
        m, n, _ = self._lps[k].dims
        x, d, yd, act = np.zeros(n), np.zeros(n), np.zeros(m), np.zeros(m)
        res = _ffi.PdhgResult()
        check(lib().sk_batch_result(self._h, k, C.byref(res), _ptr(x), _ptr(yd), _ptr(d), _ptr(act)),
              "sk_batch_result")
        return _result(res, x, yd, d, act)

    def stats(self) -> dict[str, Any]:
        s = _ffi.BatchStats()
        check(lib().sk_batch_stats_get(self._h, C.byref(s)), "sk_batch_stats_get")
        return {f: getattr(s, f) for f, _ in s._fields_}

    def time_kernel(self, kernel: int, reps: int = 20) -> tuple[float, float]:
        ms, b = C.c_double(), C.c_double()
        check(lib().sk_batch_time_kernel(self._h, kernel, reps, C.byref(ms), C.byref(b)),
              "sk_batch_time_kernel")
        return ms.value, b.value

    def probe_spmv(self, k: int, which: int, vec) -> np.ndarray:
        """The step kernels' product on instance k after a solve: which 1 gives A-hat vec
        (vec has num_cols entries), which 2 gives A-hat^T vec (num_rows entries)."""
        m, n, _ = self._lps[k].dims
        vec = np.ascontiguousarray(vec, np.float64)
        if vec.shape != ((n,) if which == 1 else (m,)):
            raise ValueError("probe_spmv: input length does not match the product")
        out = np.zeros(m if which == 1 else n)
        check(lib().sk_batch_probe_spmv(self._h, k, which, _ptr(vec), _ptr(out)),
              "sk_batch_probe_spmv")
        return out

    def probe_scaling(self, k: int) -> dict[str, np.ndarray]:
        """D_r, D_c and the scaled FP64 values (the LP's CSC order) the device set-up made."""
        m, n, z = self._lps[k].dims
        out = {"row": np.zeros(m), "column": np.zeros(n), "values": np.zeros(z)}
        check(lib().sk_batch_probe_scaling(self._h, k, _ptr(out["row"]), _ptr(out["column"]),
                                           _ptr(out["values"])), "sk_batch_probe_scaling")
        return out

    def column_order(self, k: int) -> np.ndarray:
        """The LP column each device column of instance k holds."""
        _, n, _ = self._lps[k].dims
        out = np.zeros(n, np.int32)
        check(lib().sk_batch_column_order(self._h, k, _ptr(out)), "sk_batch_column_order")
        return out

    def close(self) -> None:
        if self._h:
            lib().sk_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop on the GPU (sk_server_* of include/sk_gpu.h)."""

    def __init__(self, options: _ffi.PdhgOptions | None = None, *, slots: int = 2,
                 max_batch: int = 256, memory_budget: int = 0, keep_solutions: bool = True,
                 lanes: int | None = None, tail_fraction: float | None = None, **kw):
        self.options = options if options is not None else make_options(**kw)
        so = _ffi.ServerOptions()
        lib().sk_server_default_options(C.byref(so))
        so.slots, so.max_batch = slots, max_batch
        if lanes is not None:
            so.lanes = lanes
        if tail_fraction is not None:
            so.tail_fraction = tail_fraction
        so.memory_budget, so.keep_solutions = memory_budget, int(bool(keep_solutions))
        self.server_options = so
        h = C.c_void_p()
        check(lib().sk_server_create(C.byref(self.options), C.byref(so), C.byref(h)),
              "sk_server_create")
        self._h = h

    def now(self) -> float:
        t = C.c_double()
        check(lib().sk_server_now(self._h, C.byref(t)), "sk_server_now")
        return t.value

    def submit(self, lp: Lp, arrival: float = -1.0) -> int:
        """Hands `lp` to the server (the Lp object is spent afterwards); returns the id."""
        rid = C.c_int64()
        check(lib().sk_server_submit(self._h, lp._h, arrival, C.byref(rid)), "sk_server_submit")
        lp._h = None
        return rid.value

    def completed(self, capacity: int = 4096) -> list[int]:
        ids = (C.c_int64 * capacity)()
        cnt = C.c_int32()
        check(lib().sk_server_completed(self._h, ids, capacity, C.byref(cnt)), "sk_server_completed")
        return list(ids[:cnt.value])

    def wait(self, timeout: float = 1e9) -> bool:
        idle = C.c_int32()
        check(lib().sk_server_wait(self._h, timeout, C.byref(idle)), "sk_server_wait")
        return bool(idle.value)

    def result(self, rid: int, dims: tuple[int, int] | None = None) -> tuple[SolveResult, dict]:
        """The answer of a completed request (once). `dims` = (rows, cols) sizes the vectors;
        without it, or with keep_solutions off, only the record and scalars come back."""
        rec, res = _ffi.ServerRecord(), _ffi.PdhgResult()
        if dims is not None and self.server_options.keep_solutions:
            m, n = dims
            x, d, yd, act = np.zeros(n), np.zeros(n), np.zeros(m), np.zeros(m)
        else:
            x = d = yd = act = None
        check(lib().sk_server_result(self._h, rid, C.byref(rec), C.byref(res), _ptr(x), _ptr(yd),
                                     _ptr(d), _ptr(act)), "sk_server_result")
        record = {f: getattr(rec, f) for f, _ in rec._fields_}
        return _result(res, x, yd, d, act), record

    def stats(self) -> dict[str, Any]:
        s = _ffi.ServerStats()
        check(lib().sk_server_stats_get(self._h, C.byref(s)), "sk_server_stats_get")
        return {f: getattr(s, f) for f, _ in s._fields_}

    def close(self) -> None:
        if self._h:
            lib().sk_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
