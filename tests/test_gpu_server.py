"""The continuous-batching server (sk_server_*): every answer is the same alone and in the
mix, instances retire mid-batch, and admission respects the device-memory budget."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPTS = dict(pdhg_tolerance=1e-4, pdhg_stop_at_request=True, pdhg_absolute_primal=False)


def _drain(srv, expected, timeout=120.0):
    """Collects (result, record) for `expected` ids as they complete."""
    out = {}
    t_end = time.time() + timeout
    while len(out) < len(expected) and time.time() < t_end:
        for rid in srv.completed():
            out[rid] = srv.result(rid, expected[rid])
        time.sleep(0.002)
    assert len(out) == len(expected), f"only {len(out)} of {len(expected)} completed"
    return out


def test_served_answers_equal_solo_solves(gpu, oracle_mod):
    # C3 members of uneven size, arriving over ~0.1 s into 2 slots of at most 5 instances,
    # so they are solved in several overlapping batches of different composition
    ks = list(range(40, 56))
    srv = gpu.Server(slots=2, max_batch=5, **OPTS)
    dims, exported = {}, {}
    t0 = srv.now()
    for i, k in enumerate(ks):
        lp = gpu.Lp.generate(3, k)
        m, n, _ = lp.dims
        exported[k] = lp.export()
        rid = srv.submit(lp, arrival=t0 + 0.006 * i)
        dims[rid] = (m, n)
        time.sleep(0.006)
    got = _drain(srv, dims)
    st = srv.stats()
    srv.close()
    print("server stats", st)
    assert st["completed"] == len(ks) and st["failed"] == 0
    assert st["batches"] >= 4
    assert st["max_batch_seen"] <= 5
    ids = sorted(got)
    for rid, k in zip(ids, ks):
        res, rec = got[rid]
        assert rec["error"] == 0 and res.converged, (k, res.info["message"])
        assert rec["finished"] >= rec["admitted"] >= rec["arrival"] - 1e-9
        solo = gpu.solve(gpu.Lp.generate(3, k), **OPTS)
        # bitwise: per-instance tiles and a fixed reduction order make the batch irrelevant
        assert solo.iterations == res.iterations, k
        np.testing.assert_array_equal(solo.col_value, res.col_value)
        np.testing.assert_array_equal(solo.row_dual, res.row_dual)
        np.testing.assert_array_equal(solo.col_dual, res.col_dual)
        ev = oracle_mod.evaluate(exported[k], res.col_value, -res.row_dual)
        assert ev["worst"] <= 1e-4 * 1.000001, (k, ev["worst"])


def test_instances_retire_mid_batch(gpu):
    # one batch holding an easy and a hard LP: the easy one completes well before the batch
    # ends, while the hard one keeps running
    srv = gpu.Server(slots=1, max_batch=8, **OPTS)
    easy = [gpu.Lp.generate(3, k) for k in (60, 61, 62)]
    hard = gpu.Lp.generate(1, 1)  # 500 x 1000 at 1e-4 takes far more iterations than C3's small ones
    dims = {}
    for lp in [hard] + easy:
        m, n, _ = lp.dims
        dims[srv.submit(lp)] = (m, n)
    got = _drain(srv, dims)
    srv.close()
    recs = {rid: rec for rid, (_, rec) in got.items()}
    its = {rid: res.iterations for rid, (res, _) in got.items()}
    batches = {}
    for rid, rec in recs.items():
        batches.setdefault((rec["slot"], rec["admitted"]), []).append(rid)
    assert any(len(v) > 1 for v in batches.values())
    for members in batches.values():
        if len(members) < 2:
            continue
        last = max(members, key=lambda r: recs[r]["finished"])
        first = min(members, key=lambda r: recs[r]["finished"])
        if its[first] < its[last]:
            # the early one was read back before the batch's last instance stopped
            assert recs[first]["finished"] < recs[last]["finished"]


def test_memory_budget_admission_and_rejection(gpu):
    small = [gpu.Lp.generate(3, k) for k in (70, 71, 72, 73)]
    # a budget that holds about one C3 member at a time, so batches are single instances
    sizes = [lp.dims for lp in small]
    biggest = max(sizes, key=lambda t: t[2])
    per = 62 * biggest[2] + 110 * biggest[0] + 160 * biggest[1] + (1 << 19)
    srv = gpu.Server(slots=2, max_batch=8, memory_budget=int(1.5 * per), **OPTS)
    dims = {}
    for lp in small:
        m, n, _ = lp.dims
        dims[srv.submit(lp)] = (m, n)
    big = gpu.Lp.generate(4, 2)  # 34688 x 69376, 346880 nonzeros: far above the budget
    big_dims = big.dims
    with pytest.raises(gpu.SkError):
        srv.submit(big)
    assert big.dims == big_dims  # refused: the caller still owns a whole model
    got = _drain(srv, dims)
    st = srv.stats()
    srv.close()
    assert st["rejected"] == 1
    assert st["peak_held"] <= st["memory_budget"]
    for res, rec in got.values():
        assert rec["error"] == 0 and res.converged
    # the refused model is intact and solves on its own
    r = gpu.solve(big, **OPTS)
    assert r.converged
