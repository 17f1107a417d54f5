"""Backend selection (sk_analyze_problem, sk_select_backend): the reference's adaptive
decision (src/core/adaptive.cpp:32-232) held to the reference's own cases
(tests/unit/test_adaptive.cpp:44-293) on the CPU, with characteristics filled by hand where the
reference fills them by hand, plus the option parsing and the solve's refusal of the CPU and
hybrid backends."""
import ctypes as C

import numpy as np
import pytest

import sih26_amd as sk
from sih26_amd import _ffi


def synthetic_lp(rows, cols, per_col):
    # test_adaptive.cpp:16-36: entry (j + k) % rows of column j
    starts, idx = [0], []
    for j in range(cols):
        idx.extend(sorted({(j + k) % rows for k in range(per_col)}))
        starts.append(len(idx))
    z = len(idx)
    return sk.Lp.from_arrays(rows, cols, starts, idx, np.ones(z), np.ones(cols), np.zeros(cols),
                             np.ones(cols), np.zeros(rows), np.full(rows, 10.0))


def chars(rows, cols, nnz, gpu=False, integer=0, name=b"", memory=0, cls=b"LP"):
    c = _ffi.ProblemCharacteristics()
    c.num_rows, c.num_cols, c.num_nonzeros = rows, cols, nnz
    c.num_integer_cols, c.gpu_available = integer, int(gpu)
    c.problem_class, c.gpu_name, c.gpu_memory_bytes = cls, name, memory
    return c


GIB8 = 8 * 1024 ** 3


def test_analyzer_measures_the_model():
    c = sk.analyze_problem(synthetic_lp(10, 20, 3))
    assert (c.num_rows, c.num_cols, c.num_nonzeros, c.num_integer_cols) == (10, 20, 60, 0)
    assert c.problem_class == b"LP"
    assert c.density == pytest.approx(60 / 200) and c.sparsity_ratio == pytest.approx(1 - 60 / 200)
    assert c.estimated_matvec_flops == 120.0
    # CSC with int32 indices and FP64 values, column pointers, four FP64 vectors
    assert c.estimated_memory_bytes == 60 * 12 + 21 * 4 + 30 * 8 * 4
    assert c.gpu_available == 0 and c.gpu_device_count == 0  # no device in this container


def test_analyzer_classifies_integer_and_quadratic_models():
    lp = synthetic_lp(5, 5, 2)
    assert sk.analyze_problem(lp, integer_columns=5).problem_class == b"MILP"
    assert sk.analyze_problem(lp, quadratic=True).problem_class == b"QP"
    assert sk.analyze_problem(lp, integer_columns=1, quadratic=True).problem_class == b"MIQP"


def test_explicit_cpu_forces_the_cpu():
    c = sk.analyze_problem(synthetic_lp(2000, 2000, 30))
    d = sk.select_backend(c, sk.make_options(backend="cpu"))
    assert (d["selected"], d["requested"], d["valid"]) == ("cpu", "cpu", True)
    assert "Explicitly requested CPU" in d["reason"]


def test_explicit_gpu_fails_clearly_without_a_device():
    d = sk.select_backend(chars(5000, 5000, 100000), sk.make_options(backend="gpu"))
    assert d["requested"] == "gpu" and d["selected"] == "gpu" and not d["valid"]
    assert "no GPU device is available" in d["reason"]
    assert "[REJECTED: requested backend unavailable]" in d["explanation"]
    # the reference's spelling "cuda" selects the same mode (adaptive.hpp:51-54)
    assert not sk.select_backend(chars(5000, 5000, 100000), sk.make_options(backend="cuda"))["valid"]


def test_explicit_gpu_with_a_device_is_valid():
    d = sk.select_backend(chars(100, 100, 300, gpu=True, name=b"AMD Instinct MI355X", memory=GIB8),
                          sk.make_options(backend="gpu"))
    assert (d["selected"], d["valid"]) == ("gpu", True)
    assert "AMD Instinct MI355X" in d["explanation"]


def test_explicit_hybrid_fails_clearly_as_unsupported():
    d = sk.select_backend(chars(100, 100, 500), sk.make_options(backend="hybrid"))
    assert (d["selected"], d["requested"], d["valid"]) == ("hybrid", "hybrid", False)
    assert "unsupported in this version" in d["reason"]


def test_auto_selects_cpu_without_a_device():
    d = sk.select_backend(chars(10000, 10000, 200000), sk.make_options(backend="auto"))
    assert (d["selected"], d["requested"], d["valid"]) == ("cpu", "auto", True)
    assert "no GPU device is available" in d["reason"]


def test_auto_selects_cpu_for_a_small_problem_even_with_a_device():
    cfg = sk.adaptive_config(min_gpu_nonzeros=50000, min_gpu_dimension=1000)
    d = sk.select_backend(chars(100, 100, 300, gpu=True, memory=GIB8), None, cfg)
    assert (d["selected"], d["valid"]) == ("cpu", True)
    assert "below GPU threshold (50000 nonzeros, 1000 min dimension)" in d["reason"]


def test_auto_thresholds_are_inclusive_and_configurable():
    # has_sufficient_nonzeros is nnz >= min, has_sufficient_dimension max(rows, cols) >= min
    at = chars(10, 1000, 50000, gpu=True, memory=GIB8)
    assert sk.select_backend(at)["selected"] == "gpu"
    assert sk.select_backend(chars(10, 999, 50000, gpu=True, memory=GIB8))["selected"] == "cpu"
    assert sk.select_backend(chars(10, 1000, 49999, gpu=True, memory=GIB8))["selected"] == "cpu"
    cfg = sk.adaptive_config(adaptive_min_nnz=60000)
    assert sk.select_backend(at, None, cfg)["selected"] == "cpu"


def test_auto_selects_cpu_for_a_milp_even_with_a_device():
    d = sk.select_backend(chars(5000, 5000, 100000, gpu=True, integer=200, cls=b"MILP", memory=GIB8))
    assert (d["selected"], d["valid"]) == ("cpu", True)
    assert "MILP" in d["reason"]


def test_auto_selects_the_gpu_for_a_large_lp_with_a_device():
    cfg = sk.adaptive_config(min_gpu_nonzeros=50000, min_gpu_dimension=1000)
    d = sk.select_backend(chars(2000, 2000, 80000, gpu=True, memory=GIB8), None, cfg)
    assert (d["selected"], d["valid"]) == ("gpu", True)
    assert "exceeds GPU acceleration threshold" in d["reason"]


def test_auto_respects_the_memory_headroom():
    c = chars(200000, 500000, 2500000, gpu=True, memory=GIB8)
    c.estimated_memory_bytes = 0.81 * GIB8
    d = sk.select_backend(c)
    assert d["selected"] == "cpu" and "memory headroom" in d["reason"]
    c.estimated_memory_bytes = 0.79 * GIB8
    assert sk.select_backend(c)["selected"] == "gpu"


def test_legacy_gpu_flag_falls_back_to_the_cpu_without_a_device():
    d = sk.select_backend(chars(5000, 5000, 100000), sk.make_options(gpu=True))
    assert (d["selected"], d["requested"], d["valid"]) == ("cpu", "auto", True)
    assert "graceful fallback" in d["explanation"]
    d = sk.select_backend(chars(10, 10, 30, gpu=True, memory=GIB8), sk.make_options(gpu=True))
    assert (d["selected"], d["requested"], d["valid"]) == ("gpu", "gpu", True)
    # an explicit backend wins over the flag
    d = sk.select_backend(chars(10, 10, 30, gpu=True), sk.make_options(gpu=True, backend="cpu"))
    assert d["selected"] == "cpu"


def test_the_decision_is_deterministic_and_explained():
    c = sk.analyze_problem(synthetic_lp(50, 50, 4))
    d1, d2 = sk.select_backend(c), sk.select_backend(c)
    assert d1 == d2
    d = sk.select_backend(chars(50, 50, 150), sk.make_options(backend="auto"))
    for part in ("Backend decision: cpu", "Primary reason:", "Detailed factors:"):
        assert part in d["explanation"]


def test_adaptive_options_are_parsed_with_the_registry_bounds():
    c = sk.adaptive_config(adaptive_min_nnz="123", adaptive_min_dim="7",
                           adaptive_min_density="0.5", adaptive_explain=False)
    assert (c.min_gpu_nonzeros, c.min_gpu_dimension, c.min_gpu_density, c.explain_decision) == \
        (123, 7, 0.5, 0)
    for key, value in [("adaptive_min_nnz", "-1"), ("adaptive_min_nnz", "2000000000"),
                       ("adaptive_min_density", "2"), ("adaptive_explain", "maybe"),
                       ("adaptive_nonsense", "1")]:
        with pytest.raises(sk.SkError):
            sk.adaptive_config(**{key: value})


@pytest.mark.parametrize("backend", ["cpu", "hybrid"])
def test_solves_refuse_the_cpu_and_hybrid_backends(backend):
    lp = synthetic_lp(4, 4, 2)
    o = sk.make_options(backend=backend)
    res = _ffi.PdhgResult()
    assert sk.lib().sk_pdhg_solve(lp._h, C.byref(o), C.byref(res), None, None, None, None) == 10
    msg = sk.lib().sk_gpu_last_error().decode()
    assert ("CPU engine" in msg) if backend == "cpu" else ("unsupported" in msg)
    h = C.c_void_p()
    assert sk.lib().sk_batch_create(C.byref(o), C.byref(h)) == 10
    assert sk.lib().sk_server_create(C.byref(o), None, C.byref(h)) == 10
