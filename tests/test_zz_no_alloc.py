"""sk_gpu_dot_noalloc (the dot product whose scratch is a fixed device array of the library)
against sk_gpu_dot (library-allocated scratch), through the C ABI: the same bits on two calls
in a row, the second with a smaller input, and a clear refusal above the fixed maximum."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_N = 131072 * 2048


def device_copy(lib, host):
    ptr = C.c_void_p()
    assert lib.sk_gpu_malloc(C.byref(ptr), host.nbytes) == 0
    assert lib.sk_gpu_copy_h2d(ptr, host.ctypes.data, host.nbytes) == 0
    return ptr


def dots(lib, n, dx, dy):
    a, b = C.c_double(), C.c_double()
    assert lib.sk_gpu_dot(n, dx, dy, C.byref(a)) == 0
    assert lib.sk_gpu_dot_noalloc(n, dx, dy, C.byref(b)) == 0
    return a.value, b.value


def test_noalloc_dot_matches_the_allocating_dot(gpu):
    lib = gpu.lib()
    rng = np.random.default_rng(2026)
    x = rng.standard_normal(3_000_001)
    y = rng.standard_normal(3_000_001)
    dx, dy = device_copy(lib, x), device_copy(lib, y)
    try:
        # first call: 1,465 partials; second call smaller: 1 partial, the rest of the array
        # still holds the first call's values and must not be read
        for n in (3_000_001, 1_000):
            a, b = dots(lib, n, dx, dy)
            assert b == a, (n, a, b)
            exact = float(np.dot(x[:n], y[:n]))
            assert abs(b - exact) <= 1e-12 * float(np.sum(np.abs(x[:n] * y[:n])))
        # the empty product and the refusal above the maximum
        out = C.c_double(1.0)
        assert lib.sk_gpu_dot_noalloc(0, dx, dy, C.byref(out)) == 0 and out.value == 0.0
        assert lib.sk_gpu_dot_noalloc(MAX_N + 1, dx, dy, C.byref(out)) == 10  # invalid argument
        assert b"exceeds the maximum" in lib.sk_gpu_last_error()
    finally:
        lib.sk_gpu_free(dx)
        lib.sk_gpu_free(dy)
