"""Failure paths, forced from the host (sk_fault_inject) so that every kernel still runs to
its normal end: an allocation that fails, a launch that is refused, a non-finite instance in a
batch and in the server; and the validated primal-step dispatcher (gpu_ops_cpu.cpp:126-207)
against the oracle's CPU step."""
import ctypes as C
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL = dict(pdhg_tolerance=1e-4, pdhg_stop_at_request=True, pdhg_absolute_primal=False)


@pytest.fixture(autouse=True)
def clear_faults(gpu):
    gpu.lib().sk_fault_inject(0, 0)
    yield
    gpu.lib().sk_fault_inject(0, 0)


def p(a):
    return C.c_void_p(a.ctypes.data)


def test_validated_primal_step_is_the_oracle_step(gpu, oracle_mod):
    rng = np.random.default_rng(8)
    n = 100_003
    x, c, aty = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)
    lo = np.where(rng.random(n) < 0.3, -np.inf, -rng.random(n))
    hi = np.where(rng.random(n) < 0.3, np.inf, rng.random(n))
    hi[::97] = 1e30  # MPS-scale infinity is an absent bound (types.hpp:32-39)
    xn, xb = np.zeros(n), np.zeros(n)
    assert gpu.lib().sk_gpu_execute_pdhg_primal_step(n, 0.37, p(x), p(c), p(aty), p(lo), p(hi),
                                                     p(xn), p(xb)) == 0
    rn, rb = np.zeros(n), np.zeros(n)
    oracle_mod.lib().ok_cpu_primal_step(n, C.c_double(0.37), p(x), p(c), p(aty), p(lo), p(hi),
                                        p(rn), p(rb))
    np.testing.assert_array_equal(xn, rn)  # products rounded as on the CPU: the same bits
    np.testing.assert_array_equal(xb, rb)
    # the numerical sanity check (gpu_ops_cpu.cpp:187-193): a NaN input is reported, never
    # returned as an answer
    x[5] = np.nan
    assert gpu.lib().sk_gpu_execute_pdhg_primal_step(n, 0.37, p(x), p(c), p(aty), p(lo), p(hi),
                                                     p(xn), p(xb)) == 9
    assert b"non-finite output at column 5" in gpu.lib().sk_gpu_last_error()


def test_a_non_finite_instance_stops_alone_in_a_batch(gpu):
    lps = [gpu.Lp.generate(3, k) for k in (11, 12, 13)]
    assert gpu.lib().sk_fault_inject(3, 1) == 0  # NaN cost in instance 1 of the next upload
    b = gpu.Batch(streams=1, **REL)
    for lp in lps:
        b.add(lp)
    b.upload()
    b.solve()
    bad = b.result(1)
    assert bad.status_name == "numerical_error" and not bad.converged, bad.info["message"]
    assert "non-finite" in bad.info["message"]
    for k in (0, 2):
        r = b.result(k)
        solo = gpu.solve(lps[k], **REL)
        assert r.converged and r.iterations == solo.iterations
        np.testing.assert_array_equal(r.col_value, solo.col_value)


def test_an_allocation_failure_is_reported_and_the_next_batch_runs(gpu):
    lp = gpu.Lp.generate(3, 14)
    assert gpu.lib().sk_fault_inject(1, 5) == 0  # the fifth device allocation fails
    b = gpu.Batch(**REL)
    b.add(lp)
    with pytest.raises(gpu.SkError) as e:
        b.upload()
    assert e.value.status == 4 and "injected" in str(e.value)
    b2 = gpu.Batch(**REL)
    b2.add(lp)
    b2.upload()
    b2.solve()
    assert b2.result(0).converged


def test_a_refused_launch_fails_the_solve_and_the_next_solve_runs(gpu):
    lp = gpu.Lp.generate(3, 15)
    b = gpu.Batch(streams=1, **REL)
    b.add(lp)
    b.upload()
    assert gpu.lib().sk_fault_inject(2, 200) == 0  # a launch inside the captured PDHG block
    with pytest.raises(gpu.SkError) as e:
        b.solve()
    assert e.value.status == 5 and "injected" in str(e.value)
    b.solve()  # each solve starts from scratch and captures again
    r = b.result(0)
    assert r.converged
    assert r.iterations == gpu.solve(lp, **REL).iterations


def _drain(srv, expected, timeout=60.0):
    out, t_end = {}, time.time() + timeout
    while len(out) < len(expected) and time.time() < t_end:
        for rid in srv.completed():
            out[rid] = srv.result(rid, expected[rid])
        time.sleep(0.002)
    assert len(out) == len(expected)
    return out


def test_the_server_retires_a_failed_batch_and_serves_on(gpu):
    srv = gpu.Server(slots=1, max_batch=4, **REL)
    # a non-finite request: it completes with numerical_error, not as a failed batch
    gpu.lib().sk_fault_inject(3, 0)
    lp = gpu.Lp.generate(3, 16)
    m, n, _ = lp.dims
    got = _drain(srv, {srv.submit(lp): (m, n)})
    (res, rec), = got.values()
    assert rec["error"] == 0 and res.status_name == "numerical_error"
    # a batch whose launch is refused: its requests complete with the error ...
    gpu.lib().sk_fault_inject(2, 3)
    lp = gpu.Lp.generate(3, 17)
    m, n, _ = lp.dims
    got = _drain(srv, {srv.submit(lp): (m, n)})
    (res, rec), = got.values()
    assert rec["error"] == 5
    # ... and the next request is served normally
    lp = gpu.Lp.generate(3, 18)
    m, n, _ = lp.dims
    got = _drain(srv, {srv.submit(lp): (m, n)})
    (res, rec), = got.values()
    st = srv.stats()
    srv.close()
    assert rec["error"] == 0 and res.converged
    assert st["completed"] == 3 and st["failed"] == 1


def test_a_triangular_solve_that_gives_up_is_redone_by_rocblas(gpu):
    """The interior point's look-back triangular solves wait for earlier blocks with a bounded
    number of polls. Forced to give up at the first wait (sk_fault_inject kind 4: a bound of
    zero polls, every kernel still ends normally), each solve is redone by the rocBLAS
    fallback, and the polish ends at the same optimum as with the look-back solves."""
    lp = gpu.Lp.generate(3, 0)
    assert lp.dims[0] > 64  # more than one block, so blocks wait for each other
    first = gpu.solve(lp, **REL)
    assert first.converged
    lib = gpu.lib()
    plain = gpu.polish(lp, first)
    before = lib.sk_fault_fired(4)
    assert lib.sk_fault_inject(4, 10**6) == 0
    forced = gpu.polish(lp, first)
    lib.sk_fault_inject(0, 0)
    fired = lib.sk_fault_fired(4) - before
    assert fired >= 6 * forced.polish_iterations  # every solve of every iteration
    assert plain.status_name == forced.status_name == "optimal"
    assert abs(forced.objective - plain.objective) <= 1e-8 * (1.0 + abs(plain.objective))
    # and the look-back path works again afterwards
    again = gpu.polish(lp, first)
    assert again.status_name == "optimal"
    np.testing.assert_array_equal(again.col_value, plain.col_value)
