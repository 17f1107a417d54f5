"""GPU operators of the C ABI against the oracle's restatements of the reference's CPU
references (src/gpu/gpu_ops_cpu.cpp:21-124) and their known-answer tests
(tests/unit/test_gpu.cpp:86-177)."""
import ctypes as C

import numpy as np
import pytest

from tests.lp_cases import csc_to_dense

pytestmark = pytest.mark.gpu


class Dev:
    """Device copies of numpy float64/int32 arrays through the C ABI."""

    def __init__(self, sk):
        self.sk = sk
        self.lib = sk.lib()
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.lib.sk_gpu_malloc(C.byref(p), max(1, a.nbytes)) == 0
        if a.nbytes:
            assert self.lib.sk_gpu_copy_h2d(p, a.ctypes.data, a.nbytes) == 0
        self.ptrs.append(p)
        return p

    def empty(self, n, dtype=np.float64):
        return self.put(np.zeros(n, dtype))

    def get(self, p, n, dtype=np.float64):
        out = np.zeros(n, dtype)
        assert self.lib.sk_gpu_synchronize() == 0
        if out.nbytes:
            assert self.lib.sk_gpu_copy_d2h(out.ctypes.data, p, out.nbytes) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.lib.sk_gpu_free(p)
        self.ptrs = []


@pytest.fixture
def dev(gpu):
    d = Dev(gpu)
    yield d
    d.free()


def test_device_info(gpu):
    info = gpu.device_info(0)
    print(info)
    assert info["total_memory_bytes"] > 0
    assert info["warp_size"] == 64
    assert "gfx950" in info["gcn_arch_name"]


def test_axpy_known_answer(dev):
    # test_gpu.cpp:86-97
    x = np.array([1.0, 2.0, -3.0, 4.5])
    y = np.array([0.5, -1.0, 2.0, 0.0])
    dx, dy = dev.put(x), dev.put(y)
    assert dev.lib.sk_gpu_axpy(4, 2.0, dx, dy) == 0
    np.testing.assert_array_equal(dev.get(dy, 4), [2.5, 3.0, -4.0, 9.0])


def test_dot_known_answer_and_large(dev, oracle_mod):
    # test_gpu.cpp:99-105: 10.5
    x = np.array([1.0, 2.0, -3.0, 4.0])
    y = np.array([2.0, -1.0, 0.5, 3.0])
    out = C.c_double()
    assert dev.lib.sk_gpu_dot(4, dev.put(x), dev.put(y), C.byref(out)) == 0
    assert out.value == 10.5
    rng = np.random.default_rng(3)
    n = 1_000_003
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    assert dev.lib.sk_gpu_dot(n, dev.put(x), dev.put(y), C.byref(out)) == 0
    exact = float(np.dot(x.astype(np.longdouble), y.astype(np.longdouble)))
    # FP64 summation error bound: n * eps * sum |x_i y_i|
    assert abs(out.value - exact) <= n * 2.2e-16 * float(np.sum(np.abs(x * y)))
    # deterministic: the same bits twice
    out2 = C.c_double()
    assert dev.lib.sk_gpu_dot(n, dev.ptrs[-2], dev.ptrs[-1], C.byref(out2)) == 0
    assert out2.value == out.value


def test_primal_step_known_answer(dev, oracle_mod):
    # test_gpu.cpp:107-140
    inf = float("inf")
    x = np.array([2.0, 0.0, -4.0, 0.5])
    cost = np.array([1.0, -2.0, 2.0, 3.0])
    aty = np.array([0.0, 1.0, 0.0, 0.0])
    lo = np.array([-1.0, 0.0, -inf, 0.0])
    hi = np.array([1.0, 2.0, inf, inf])
    dn, de = dev.empty(4), dev.empty(4)
    assert dev.lib.sk_gpu_pdhg_primal_step(4, 0.5, dev.put(x), dev.put(cost), dev.put(aty),
                                           dev.put(lo), dev.put(hi), dn, de) == 0
    np.testing.assert_array_equal(dev.get(dn, 4), [1.0, 0.5, -5.0, 0.0])
    np.testing.assert_array_equal(dev.get(de, 4), [0.0, 1.0, -6.0, -0.5])


def test_primal_step_matches_oracle_bitwise(dev, oracle_mod):
    rng = np.random.default_rng(11)
    n = 100_000
    x, c, aty = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)
    lo = np.where(rng.random(n) < 0.2, -np.inf, -rng.random(n))
    hi = np.where(rng.random(n) < 0.2, 1e30, rng.random(n))
    dn, de = dev.empty(n), dev.empty(n)
    assert dev.lib.sk_gpu_pdhg_primal_step(n, 0.37, dev.put(x), dev.put(c), dev.put(aty),
                                           dev.put(lo), dev.put(hi), dn, de) == 0
    gx, ge = dev.get(dn, n), dev.get(de, n)
    rx, re = np.zeros(n), np.zeros(n)
    L = oracle_mod.lib()
    L.ok_cpu_primal_step(n, 0.37, x.ctypes.data, c.ctypes.data, aty.ctypes.data, lo.ctypes.data,
                         hi.ctypes.data, rx.ctypes.data, re.ctypes.data)
    # x - tau*g may fuse into one FMA on the GPU: allow one rounding of difference
    np.testing.assert_allclose(gx, rx, rtol=4e-16, atol=4e-16 * np.abs(0.37 * (c + aty)).max())
    np.testing.assert_allclose(ge, 2 * gx - x, rtol=0, atol=0)


def _csc_2x3():
    # test_gpu.cpp:142-177
    return (np.array([0, 1, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32),
            np.array([1.0, 3.0, 2.0, 4.0]))


def test_spmv_csc_both_orientations_known_answer(dev):
    cs, ri, v = _csc_2x3()
    dcs, dri, dv = dev.put(cs), dev.put(ri), dev.put(v)
    y = dev.empty(2)
    assert dev.lib.sk_gpu_spmv_csc(2, 3, dcs, dri, dv, dev.put(np.array([1.0, 2.0, 3.0])), y, 1.0, 0) == 0
    np.testing.assert_array_equal(dev.get(y, 2), [7.0, 18.0])
    yt = dev.empty(3)
    assert dev.lib.sk_gpu_spmv_csc(2, 3, dcs, dri, dv, dev.put(np.array([2.0, 5.0])), yt, 1.0, 1) == 0
    np.testing.assert_array_equal(dev.get(yt, 3), [2.0, 15.0, 24.0])


def test_spmv_csr_matches_dense(dev, gpu):
    lp = gpu.Lp.generate(3, 7).export()
    a = csc_to_dense(lp)
    m, n = a.shape
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n)
    # CSR of A = CSC of A^T
    rows = [np.nonzero(a[i])[0] for i in range(m)]
    ptr = np.zeros(m + 1, np.int32)
    ptr[1:] = np.cumsum([len(r) for r in rows])
    idx = np.concatenate(rows).astype(np.int32)
    val = np.concatenate([a[i, r] for i, r in enumerate(rows)])
    y = dev.empty(m)
    assert dev.lib.sk_gpu_spmv_csr(m, n, dev.put(ptr), dev.put(idx), dev.put(val), dev.put(x), y, 1.5) == 0
    got = dev.get(y, m)
    exact = 1.5 * (a @ x)
    bound = 1.5 * (np.abs(a) @ np.abs(x)) * 64 * 2.2e-16
    assert np.all(np.abs(got - exact) <= bound + 1e-300)


def test_compare_vectors_semantics(gpu):
    # test_gpu.cpp:183-216
    lib = gpu.lib()
    from sih26_amd import _ffi

    def cmp(a, b, tol):
        a, b = np.asarray(a, float), np.asarray(b, float)
        r = _ffi.Comparison()
        assert lib.sk_gpu_compare_vectors(len(a), a.ctypes.data, b.ctypes.data, tol, C.byref(r)) == 0
        return r

    r = cmp([1.0, 2.0, 3.0], [1.0, 2.00000001, 2.99999999], 1e-7)
    assert r.passed and not r.contains_nan_or_inf and r.max_abs_diff <= 1e-7
    r = cmp([1.0, 2.0, 3.0], [1.0, 2.05, 3.0], 1e-4)
    assert not r.passed and abs(r.max_abs_diff - 0.05) < 1e-6
    r = cmp([1.0, float("nan"), 3.0], [1.0, 2.0, 3.0], 1e-4)
    assert not r.passed and r.contains_nan_or_inf
    r = cmp([1.0, float("inf"), 3.0], [1.0, 2.0, 3.0], 1e-4)
    assert not r.passed and r.contains_nan_or_inf
