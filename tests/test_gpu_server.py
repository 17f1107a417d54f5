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
        # bitwise: per-instance tiles and a fixed reduction order make the batch irrelevant at
        # one tile capacity (these batches and the solo solves are all below a quarter of a
        # round of resident workgroups, so all run at 1,024 nonzeros per tile)
        assert solo.iterations == res.iterations, k
        np.testing.assert_array_equal(solo.col_value, res.col_value)
        np.testing.assert_array_equal(solo.row_dual, res.row_dual)
        np.testing.assert_array_equal(solo.col_dual, res.col_dual)
        ev = oracle_mod.evaluate(exported[k], res.col_value, -res.row_dual)
        assert ev["worst"] <= 1e-4 * 1.000001, (k, ev["worst"])


def test_answers_at_different_default_capacities_agree_to_the_tolerance(gpu, oracle_mod):
    """The default tile capacity follows a batch's size: 64 C3 members as one launch (streams=1)
    fill more than a quarter of a round of resident workgroups and run at 1,536, each alone at
    1,024. The per-tile partials then group differently, so answers need not agree bit for
    bit; each meets the request (oracle relative KKT <= 1e-4) in both, and the objectives agree
    to 1e-3 of the objectives' scale."""
    ks = list(range(64))
    lps = [gpu.Lp.generate(3, k) for k in ks]
    b = gpu.Batch(streams=1, **OPTS)
    for lp in lps:
        b.add(lp)
    b.upload()
    b.solve()
    assert b.stats()["tile_nnz"] == 1536
    for k in (0, 21, 42, 63):
        d = lps[k].export()
        mixed = b.result(k)
        one = gpu.Batch(streams=1, **OPTS)
        one.add(lps[k])
        one.upload()
        one.solve()
        assert one.stats()["tile_nnz"] == 1024
        alone = one.result(0)
        for r in (mixed, alone):
            assert r.converged, (k, r.info["message"])
            ev = oracle_mod.evaluate(d, r.col_value, -d["sense"] * r.row_dual)
            assert ev["worst"] <= 1e-4 * 1.000001, (k, ev["worst"])
        scale = 1.0 + abs(alone.primal_objective) + abs(alone.dual_objective)
        assert abs(mixed.objective - alone.objective) <= 1e-3 * scale, (k, mixed.objective, alone.objective)


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


def _tail_scenario(gpu, lanes):
    """One slot of at most 4 instances. A first request keeps the slot busy while the first
    wave (one 40,000 x 80,000 request and three small ones) and the second wave (four small
    ones) queue behind it, so the first wave is one batch and the second comes next."""
    big = gpu.Lp.generate_params(**dict(gpu.Lp.preset(4, 0), num_rows=40000, num_cols=80000,
                                        num_equality_rows=24000))
    lps = [gpu.Lp.generate(3, 60)] + [big] + [gpu.Lp.generate(3, k) for k in (61, 62, 63)] + \
        [gpu.Lp.generate(3, k) for k in (64, 65, 66, 67)]
    dims = [lp.dims[:2] for lp in lps]
    srv = gpu.Server(slots=1, max_batch=4, lanes=lanes, **OPTS)
    ids = [srv.submit(lps[0])]
    t_end = time.time() + 30.0
    while srv.stats()["batches"] < 1 and time.time() < t_end:  # the first request is running
        time.sleep(0.0002)
    ids += [srv.submit(lp) for lp in lps[1:]]
    got = _drain(srv, dict(zip(ids, dims)), timeout=300.0)
    st = srv.stats()
    srv.close()
    return [got[i] for i in ids], st


def test_a_batch_in_its_tail_lets_the_next_one_start(gpu):
    """With two lanes, the second wave is admitted once the first wave's batch is down to its
    last instance (the large one), long before that instance finishes; with one lane it waits
    for the whole batch. Every answer is the same either way, bit for bit."""
    two, st2 = _tail_scenario(gpu, 2)
    one, st1 = _tail_scenario(gpu, 1)
    for got, st in ((two, st2), (one, st1)):
        assert st["completed"] == 9 and st["failed"] == 0
        assert all(rec["error"] == 0 and res.converged for res, rec in got)
        assert got[1][1]["admitted"] == got[4][1]["admitted"]  # the first wave is one batch
    big_done_2, big_done_1 = two[1][1]["finished"], one[1][1]["finished"]
    wave2_in_2 = min(rec["admitted"] for _, rec in two[5:])
    wave2_in_1 = min(rec["admitted"] for _, rec in one[5:])
    small_done_2 = max(rec["finished"] for _, rec in two[2:5])
    print(f"two lanes: second wave admitted {1e3 * (wave2_in_2 - small_done_2):.2f} ms after the "
          f"first wave's small ones, {1e3 * (big_done_2 - wave2_in_2):.1f} ms before the large one "
          f"finished; one lane: {1e3 * (wave2_in_1 - big_done_1):.2f} ms after it")
    assert wave2_in_2 < big_done_2
    assert wave2_in_1 >= big_done_1
    for (a, _), (b, _) in zip(two, one):
        assert a.iterations == b.iterations
        np.testing.assert_array_equal(a.col_value, b.col_value)
        np.testing.assert_array_equal(a.row_dual, b.row_dual)
