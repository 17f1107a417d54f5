"""The CPU-only build (sih26_amd/lib/libsk_gpu_stub.so): the counterpart of the reference's
no-CUDA build, which links src/gpu/gpu_runtime_stub.cpp:11-82 instead of the device code. It
exports every entry point of include/sk_gpu.h; the host parts (models, generator, MPS reader
and writer, instance cache, option parser, backend selection) work; every device entry point
returns SK_GPU_UNAVAILABLE and the device queries report none."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from sih26_amd import _ffi
from tests.test_cabi_host import declared_functions

STUB = Path(_ffi.LIB_PATH).parent / "libsk_gpu_stub.so"
UNAVAILABLE = 1


@pytest.fixture(scope="module")
def stub():
    if not STUB.exists():
        pytest.fail(f"{STUB} is not built (python sih26_amd/build.py)")
    return _ffi.load(STUB)


def test_stub_exports_every_declared_function_and_links_no_device_runtime(stub):
    lib = C.CDLL(str(STUB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    import subprocess
    deps = subprocess.run(["ldd", str(STUB)], capture_output=True, text=True).stdout
    assert "amdhip" not in deps and "rocblas" not in deps


def test_device_entry_points_report_no_device(stub):
    assert stub.sk_gpu_is_available() == 0 and stub.sk_gpu_device_count() == 0
    ptr = C.c_void_p(1)
    assert stub.sk_gpu_malloc(C.byref(ptr), 64) == UNAVAILABLE and not ptr.value
    assert b"CPU-only build" in stub.sk_gpu_last_error()
    out = C.c_double(7.0)
    assert stub.sk_gpu_dot(4, None, None, C.byref(out)) == UNAVAILABLE and out.value == 0.0
    assert stub.sk_gpu_dot_noalloc(4, None, None, C.byref(out)) == UNAVAILABLE
    assert stub.sk_gpu_status_string(UNAVAILABLE) == b"hip unavailable"
    info = _ffi.DeviceInfo()
    assert stub.sk_gpu_get_device_info(0, C.byref(info)) == UNAVAILABLE


def test_host_parts_work_and_solves_report_unavailable(stub, tmp_path):
    lp = C.c_void_p()
    assert stub.sk_lp_generate(1, 1, C.byref(lp)) == 0
    m, n, z = C.c_int32(), C.c_int32(), C.c_int64()
    assert stub.sk_lp_dims(lp, C.byref(m), C.byref(n), C.byref(z)) == 0
    assert (m.value, n.value, z.value) == (500, 1000, 5000)
    path = str(tmp_path / "c1.mps").encode()
    assert stub.sk_lp_write_mps(lp, path) == 0
    back, k = C.c_void_p(), C.c_int32()
    assert stub.sk_lp_read_mps(path, C.byref(back), C.byref(k)) == 0
    vals = np.zeros(5000)
    assert stub.sk_lp_export(back, None, None, None, None, C.c_void_p(vals.ctypes.data), None,
                             None, None, None, None) == 0
    assert vals.min() != 0.0
    o = _ffi.PdhgOptions()
    stub.sk_pdhg_default_options(C.byref(o))
    assert stub.sk_pdhg_set_option(C.byref(o), b"pdhg_tolerance", b"1e-6") == 0
    res = _ffi.PdhgResult()
    assert stub.sk_pdhg_solve(lp, C.byref(o), C.byref(res), None, None, None, None) == UNAVAILABLE
    assert b"no HIP device code" in stub.sk_gpu_last_error()
    # the selection of this build: no device, so auto picks the CPU and an explicit GPU request
    # is invalid (adaptive.cpp:144-158, 176-184)
    ch, d = _ffi.ProblemCharacteristics(), _ffi.BackendDecision()
    assert stub.sk_analyze_problem(lp, 0, 0, C.byref(ch)) == 0 and ch.gpu_available == 0
    assert stub.sk_select_backend(C.byref(ch), C.byref(o), None, C.byref(d)) == 0
    assert d.selected_backend == 1 and d.is_valid == 1
    assert stub.sk_pdhg_set_option(C.byref(o), b"backend", b"gpu") == 0
    assert stub.sk_select_backend(C.byref(ch), C.byref(o), None, C.byref(d)) == 0
    assert d.selected_backend == 2 and d.is_valid == 0
    for h in (lp, back):
        stub.sk_lp_free(h)
