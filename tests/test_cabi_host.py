"""Host-side checks of the C ABI (no GPU work): the library loads and exports every function
include/sk_gpu.h declares; options, model validation and the seeded generator behave."""
import ctypes as C
import hashlib
import re
from pathlib import Path

import numpy as np
import pytest

import sih26_amd as sk
from sih26_amd import _ffi
from tests.lp_cases import dense_to_lp

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "sk_gpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = C.CDLL(str(_ffi.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python bindings declare a signature for each
    assert set(names) <= set(_ffi.SIGNATURES), sorted(set(names) - set(_ffi.SIGNATURES))


def test_status_strings_are_pinned():
    # the reference pins its status strings (tests/unit/test_gpu.cpp:61-69)
    s = lambda k: sk.lib().sk_gpu_status_string(k).decode()  # noqa: E731
    assert s(0) == "success"
    assert s(4) == "out of device memory"
    assert s(5) == "kernel launch failed"
    assert s(9) == "non-finite numerical output detected"
    assert s(10) == "invalid argument"
    assert s(99) == "unknown status"


def test_option_registry_semantics():
    o = sk.make_options()
    assert o.tolerance == 1e-4 and o.iteration_limit == -1 and o.restart == 1
    assert o.stop_at_request == 0 and o.absolute_primal == 1 and o.precision == 0
    o = sk.make_options(pdhg_tolerance="1e-6", iteration_limit=50, pdhg_precision="FP64",
                        pdhg_restart=False, algorithm="PDHG", backend="gpu")
    assert o.tolerance == 1e-6 and o.iteration_limit == 50 and o.precision == 1 and o.restart == 0
    for key, value in [("no_such_option", "1"), ("pdhg_tolerance", "abc"),
                       ("pdhg_tolerance", "0.5"), ("algorithm", "simplex"),
                       ("backend", "quantum"), ("pdhg_polish", "maybe"), ("iteration_limit", "-2"),
                       ("pdhg_precision", "fp16")]:
        with pytest.raises(sk.SkError) as e:
            sk.make_options(**{key: value})
        assert e.value.status == 10
    # pdhg_polish is validated and accepted; the polish itself is sk_pdhg_polish
    sk.make_options(pdhg_polish="true")
    sk.make_options(pdhg_polish="false")


def test_model_validation_rejects_malformed_models():
    good = dense_to_lp([[1.0, 1.0]], [2.0], [np.inf], [1.0, 1.0])

    def create(**change):
        d = dict(good, **change)
        return sk.Lp.from_arrays(d["num_rows"], d["num_cols"], d["col_starts"], d["row_indices"],
                                 d["values"], d["cost"], d["col_lower"], d["col_upper"],
                                 d["row_lower"], d["row_upper"])

    assert create().dims == (1, 2, 2)
    bad = [dict(col_lower=np.array([0.0, np.nan])),
           dict(col_lower=np.array([3.0, 0.0]), col_upper=np.array([1.0, np.inf])),
           dict(row_lower=np.array([5.0]), row_upper=np.array([1.0])),
           dict(values=np.array([1.0, np.inf])),
           dict(cost=np.array([np.nan, 1.0])),
           dict(row_indices=np.array([0, 3], np.int32))]
    for change in bad:
        with pytest.raises(sk.SkError) as e:
            create(**change)
        assert e.value.status == 10
    # rows must be strictly ascending inside a column (frozen CSC, sparse.cpp:82-125)
    two_rows = dense_to_lp([[1.0], [1.0]], [0.0, 0.0], [1.0, 1.0], [1.0])
    with pytest.raises(sk.SkError):
        sk.Lp.from_arrays(2, 1, two_rows["col_starts"], np.array([1, 0], np.int32),
                          two_rows["values"], two_rows["cost"], two_rows["col_lower"],
                          two_rows["col_upper"], two_rows["row_lower"], two_rows["row_upper"])


def test_lp_round_trips_through_the_library():
    d = dense_to_lp([[1.0, 0.0, 2.0], [0.0, 3.0, 4.0]], [1.0, -np.inf], [1.0, 5.0],
                    [1.0, -1.0, 0.5], [np.inf, 2.0, 1e30], sense=-1, offset=3.5)
    lp = sk.Lp.from_arrays(d["num_rows"], d["num_cols"], d["col_starts"], d["row_indices"],
                           d["values"], d["cost"], d["col_lower"], d["col_upper"],
                           d["row_lower"], d["row_upper"], -1, 3.5)
    e = lp.export()
    assert e["sense"] == -1 and e["offset"] == 3.5
    for k in ("col_starts", "row_indices", "values", "cost", "col_lower", "col_upper",
              "row_lower", "row_upper"):
        np.testing.assert_array_equal(e[k], d[k])


def digest(lp):
    e = lp.export()
    h = hashlib.sha256()
    for k in ("col_starts", "row_indices", "values", "cost", "col_lower", "col_upper",
              "row_lower", "row_upper"):
        h.update(np.ascontiguousarray(e[k]).tobytes())
    return h.hexdigest()


def test_generator_is_a_pure_function_of_its_seed():
    assert digest(sk.Lp.generate(1, 1)) == digest(sk.Lp.generate(1, 1))
    assert digest(sk.Lp.generate(1, 1)) != digest(sk.Lp.generate(1, 2))
    assert digest(sk.Lp.generate(3, 7)) == digest(sk.Lp.generate(3, 7))


def test_blending_lp_shape_and_feasibility_by_construction():
    e = sk.Lp.generate(1, 1).export()
    m, n = e["num_rows"], e["num_cols"]
    assert (m, n) == (500, 1000)
    cs, ri, v = e["col_starts"], e["row_indices"], e["values"]
    assert np.all(np.diff(cs) == 5) and cs[-1] == 5000
    for j in range(n):
        rows = ri[cs[j]:cs[j + 1]]
        assert np.all(np.diff(rows) > 0)
    assert np.all(np.bincount(ri, minlength=m) > 0)  # no empty row
    mag = np.abs(v)
    assert mag.min() >= 0.1 and mag.max() <= 2.0
    assert 0.15 < np.mean(v < 0) < 0.25
    assert np.all(e["row_lower"][:300] == e["row_upper"][:300])
    assert np.all(np.isinf(e["row_lower"][300:]))
    assert np.all(e["col_lower"] == 0) and np.all((e["col_upper"] >= 1) & (e["col_upper"] <= 100))
    assert np.all(np.abs(e["cost"]) <= 1)


def test_batch_member_shapes():
    for k in (0, 1, 255):
        p = sk.Lp.preset(3, k)
        m, n = p["num_rows"], p["num_cols"]
        assert 200 <= m <= 5000 and 2 * m - 1 <= n <= 3 * m + 1
        assert p["degree_lo"] == 4 and p["degree_hi"] == 8 and p["seed"] == 1000 + k
    assert Path  # keep import used


def test_power_law_preset_hits_exact_nnz():
    # a smaller instance of the power-law family: trimmed/padded to exactly target_nnz
    lp = sk.Lp.generate_params(**dict(sk.Lp.preset(2, 11), num_rows=2000, num_cols=5000,
                                      num_equality_rows=1200, target_nnz=25000))
    m, n, z = lp.dims
    assert (m, n, z) == (2000, 5000, 25000)
    deg = np.diff(lp.export()["col_starts"])
    assert deg.max() <= 2000 and deg.min() >= 1


def test_server_options_and_missing_device():
    so = _ffi.ServerOptions()
    sk.lib().sk_server_default_options(C.byref(so))
    assert (so.slots, so.max_batch, so.memory_budget, so.keep_solutions) == (2, 256, 0, 1)
    assert (so.lanes, so.tail_fraction) == (1, 0.25)
    if sk.gpu_available():
        pytest.skip("a device is visible; the missing-device path is for CPU-only hosts")
    h = C.c_void_p()
    st = sk.lib().sk_server_create(None, C.byref(so), C.byref(h))
    assert st == 2 and not h.value  # SK_GPU_NO_DEVICE, nothing created
    bad = _ffi.ServerOptions()
    sk.lib().sk_server_default_options(C.byref(bad))
    bad.slots = 0
    assert sk.lib().sk_server_create(None, C.byref(bad), C.byref(h)) == 10  # invalid argument
    for field, value in (("lanes", 0), ("lanes", 9), ("tail_fraction", 1.5), ("tail_fraction", float("nan"))):
        sk.lib().sk_server_default_options(C.byref(bad))
        setattr(bad, field, value)
        assert sk.lib().sk_server_create(None, C.byref(bad), C.byref(h)) == 10, field
    # null handles are argument errors, not crashes
    assert sk.lib().sk_server_submit(None, None, -1.0, None) == 10


def test_ipm_options_and_argument_checks():
    o = _ffi.IpmOptions()
    sk.lib().sk_ipm_default_options(C.byref(o))
    # the reference's polish defaults (src/util/options.cpp:355-393, 183)
    assert (o.iteration_limit, o.start_margin, o.max_seconds, o.max_factor_nonzeros, o.scaling) == \
        (50, 0.1, 30.0, 50_000_000, 1)
    nulls = [None] * 11
    assert sk.lib().sk_pdhg_polish(*nulls) == 10
    assert sk.lib().sk_ipm_solve(None, None, None, None, None, None, None, None, None, None) == 10
    lp = sk.Lp.generate(1, 1)
    res = _ffi.PdhgResult()
    for field, value in (("start_margin", 0.0), ("max_seconds", -1.0)):
        bad = _ffi.IpmOptions()
        sk.lib().sk_ipm_default_options(C.byref(bad))
        setattr(bad, field, value)
        st = sk.lib().sk_ipm_solve(lp._h, C.byref(bad), None, None, None, C.byref(res),
                                   None, None, None, None)
        assert st == 10, field  # rejected before any device work


def test_fault_injection_arguments_are_checked():
    lib = sk.lib()
    assert lib.sk_fault_inject(1, 3) == 0 and lib.sk_fault_inject(2, 1) == 0
    assert lib.sk_fault_inject(3, 0) == 0 and lib.sk_fault_inject(4, 2) == 0
    assert lib.sk_fault_inject(0, 0) == 0
    for kind, n in ((1, 0), (2, -1), (3, -2), (4, 0), (5, 1), (-1, 0)):
        assert lib.sk_fault_inject(kind, n) == 10
    # nothing fired: no device work ran
    assert [lib.sk_fault_fired(k) for k in (1, 2, 3, 4)] == [0, 0, 0, 0]
    assert lib.sk_fault_fired(0) == 0 and lib.sk_fault_fired(9) == 0


def test_validated_primal_step_without_a_device_reports_unavailable():
    # no device in this container: the dispatcher returns the reference's kCudaUnavailable
    # class (SK_GPU_UNAVAILABLE) and the caller runs its CPU step (gpu_ops_cpu.cpp:131-134)
    n = 4
    a = [np.ones(n) for _ in range(7)]
    ptrs = [C.c_void_p(v.ctypes.data) for v in a]
    lib = sk.lib()
    assert lib.sk_gpu_execute_pdhg_primal_step(0, 1.0, *ptrs[:7]) == 0  # empty: success
    if lib.sk_gpu_is_available():
        pytest.skip("a device is present")
    assert lib.sk_gpu_execute_pdhg_primal_step(n, 1.0, *ptrs[:7]) == 1
    assert lib.sk_gpu_execute_pdhg_primal_step(n, 1.0, None, *ptrs[1:7]) == 10
