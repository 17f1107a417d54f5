"""Benchmark: LPs solved/s and PDHG iterations/s at 1e-4 relative KKT on one MI355X.

Default workload: the 256-LP batch of uneven blending LPs (rows U{200..5000}, columns 2-3x
rows, 4-8 nonzeros per column, seeds 1000+k), solved as one block-diagonal batch by the
FP32/FP64-split engine with the relative-KKT stop (pdhg_stop_at_request, no absolute-primal
clause). One "step" is one complete solve of the batch from inputs resident in HBM: device
scaling, power iteration, the PDHG loop until every instance stops, and the final report.

Prints one JSON line. Extra fields:
  roofline      the dominant step kernel, timed with HIP events on the engine's stream
  cpu_baseline  the FP64 CPU oracle on a bounded sample of the same instances, same stop rule
  fp64_single   the plain FP64 engine solving instances one at a time (the x20 denominator)
  serving       the request loop (sk_server_*) on a Poisson stream of blending LPs (rows
                log-uniform on [500, 50000], 2 columns per row, 5 nonzeros per column, seeds
                100000+i; arrivals at 200/s from seed 42): p50/p99 of completion - arrival

`--workload c4` makes the serving stream the main line (10,000 requests by default).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "c3": "256 blending LPs, rows U{200..5000}, cols 2-3x rows, 4-8 nnz/col, block-diagonal batch",
    "c1": "one 500x1000 blending LP, 5 nnz/col",
    "c2": "one 200k x 500k power-law LP, 2.5M nnz, |a| log-uniform [1e-3, 1e3]",
    "c5": "one 2M x 5M power-law LP, 25M nnz",
    "c4": "request stream at 200/s of blending LPs, rows log-uniform [500, 50000], n = 2m, "
          "5 nnz/col, continuous batching",
}
KERNEL_NAMES = {0: "k_step_primal", 1: "k_step_dual", 2: "k_step_aty"}


def solve_options(sk, precision: str = "fp32", tolerance: float = 1e-4, **extra):
    return sk.make_options(pdhg_tolerance=tolerance, pdhg_stop_at_request=True,
                           pdhg_absolute_primal=False, pdhg_precision=precision,
                           compute_quality=0, **extra)


def instances(sk, workload: str, count: int):
    if workload == "c3":
        return [sk.Lp.generate(3, k) for k in range(count)]
    cfg = {"c1": 1, "c2": 2, "c5": 5}[workload]
    seed = {"c1": 1, "c2": 11, "c5": 5}[workload]
    return [sk.Lp.generate(cfg, seed)]


def host_cpu() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    """The host threads this process may use, at most 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(lps, budget_s: float) -> dict:
    """The FP64 CPU oracle (restatement of src/pdhg/pdhg.cpp), same stop rule, on instances
    0, 1, 2, ... until `budget_s` of CPU time is used: one thread, then the same instances
    with the A^T y gather on every host thread (the reference's OpenMP loop,
    sparse.cpp:216-220). Rates with and without the set-up (scaling + power iteration).
    A single-LP workload is run under a time limit of `budget_s` and reported in PDHG
    iterations/s of the loop (set-up excluded): the oracle does not finish it in minutes."""
    from oracle import oracle
    oracle.build(ref=False, arch_flags="-march=native")
    oracle.lib()
    exported = []
    if len(lps) == 1:
        return cpu_baseline_single(oracle, lps[0].export(), budget_s)

    def run(sample):
        done, iters, setup, t0 = 0, 0, 0.0, time.perf_counter()
        for lp in sample:
            r = oracle.solve_pdhg(lp, tolerance=1e-4, stop_at_request=True,
                                  absolute_primal=False)
            done += 1
            iters += r["iterations"]
            setup += r["setup_seconds"]
            if sample is not exported:
                exported.append(lp)
            if time.perf_counter() - t0 > budget_s:
                break
        return done, iters, setup, time.perf_counter() - t0

    oracle.set_threads(1)
    done, iters, setup, dt = run(lp.export() for lp in lps)
    threads = host_threads()
    oracle.set_threads(threads)
    done_t, iters_t, setup_t, dt_t = run(exported)
    oracle.set_threads(1)
    parallel = cpu_instance_parallel(exported[:done], threads)
    return {"value": done / dt, "unit": "LPs/s", "cores": 1, "kind": "port",
            "sample": f"first {done} of the {len(lps)} instances, FP64 oracle, 1 thread, "
                      f"{dt:.1f} s, {iters / dt:.0f} PDHG iterations/s",
            "iters_per_s": iters / dt,
            "value_without_setup": done / (dt - setup), "setup_s": setup,
            "host_cpu": host_cpu(),
            "all_threads": {"value": done_t / dt_t, "cores": threads,
                            "iters_per_s": iters_t / dt_t,
                            "value_without_setup": done_t / (dt_t - setup_t),
                            "sample": f"the same {done_t} instances, A^T y gather on {threads} "
                                      f"threads, {dt_t:.1f} s"},
            "instance_parallel": parallel}


_WORKER = r'''
import json, sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
from oracle import oracle
oracle.lib()
t0 = time.perf_counter()
iters = 0
for path in sys.argv[2:]:
    z = np.load(path)
    lp = {k: z[k] for k in z.files}
    for k in ("num_rows", "num_cols", "sense"):
        lp[k] = int(lp[k])
    lp["offset"] = float(lp["offset"])
    iters += oracle.solve_pdhg(lp, tolerance=1e-4, stop_at_request=True,
                               absolute_primal=False)["iterations"]
print(json.dumps({"iterations": iters, "seconds": time.perf_counter() - t0}))
'''


def cpu_instance_parallel(sample: list, procs: int) -> dict:
    """The same instances spread over `procs` child processes (one oracle thread each), the
    way a batch of independent LPs uses a multi-core host."""
    import subprocess
    import tempfile
    import numpy as np
    with tempfile.TemporaryDirectory() as tmp:
        files = []
        for k, lp in enumerate(sample):
            f = Path(tmp) / f"lp{k}.npz"
            np.savez(f, **{key: np.asarray(v) for key, v in lp.items()})
            files.append(str(f))
        t0 = time.perf_counter()
        children = [subprocess.Popen([sys.executable, "-c", _WORKER, str(ROOT)] + files[r::procs],
                                     stdout=subprocess.PIPE, text=True)
                    for r in range(procs) if files[r::procs]]
        outs = [c.communicate()[0] for c in children]
        wall = time.perf_counter() - t0
        if any(c.returncode for c in children):
            return {"error": "a worker failed"}
    iters = sum(json.loads(o)["iterations"] for o in outs)
    return {"value": len(sample) / wall, "cores": len(children), "iters_per_s": iters / wall,
            "sample": f"the same {len(sample)} instances over {len(children)} processes, one "
                      f"oracle thread each, {wall:.1f} s wall (process start included)"}


def cpu_baseline_single(oracle, lp: dict, budget_s: float) -> dict:
    legs = []
    for threads in (1, host_threads()):
        oracle.set_threads(threads)
        solves, iters, setup, t0 = 0, 0, 0.0, time.perf_counter()
        while True:
            r = oracle.solve_pdhg(lp, tolerance=1e-4, stop_at_request=True,
                                  absolute_primal=False, time_limit=budget_s)
            solves += 1
            iters += r["iterations"]
            setup += r["setup_seconds"]
            if r["status"] == 7 or time.perf_counter() - t0 > budget_s:
                break
        legs.append((threads, solves, iters, setup, time.perf_counter() - t0, r["status"]))
    oracle.set_threads(1)
    (_, n1, it1, su1, dt1, st1), (threads, nt, itt, sut, dtt, stt) = legs
    if st1 != 7 and stt != 7:  # the LP finishes within the budget: LPs/s as for a batch
        return {"value": n1 / dt1, "unit": "LPs/s", "cores": 1, "kind": "port",
                "sample": f"the LP solved {n1} times ({it1 // n1} iterations each), FP64 "
                          f"oracle, 1 thread, {dt1:.1f} s",
                "iters_per_s": it1 / dt1, "value_without_setup": n1 / (dt1 - su1),
                "setup_s": su1 / n1, "host_cpu": host_cpu(),
                "all_threads": {"value": nt / dtt, "cores": threads, "iters_per_s": itt / dtt,
                                "value_without_setup": nt / (dtt - sut),
                                "sample": f"solved {nt} times, A^T y gather on {threads} "
                                          "threads"}}
    return {"value": it1 / (dt1 - su1), "unit": "PDHG iterations/s", "cores": 1, "kind": "port",
            "sample": f"the LP's first {it1} iterations (time limit {budget_s:.0f} s, status "
                      f"{st1}), FP64 oracle, 1 thread; set-up {su1:.1f} s excluded",
            "setup_s": su1, "host_cpu": host_cpu(),
            "all_threads": {"value": itt / (dtt - sut), "cores": threads, "setup_s": sut,
                            "sample": f"first {itt} iterations, A^T y gather on {threads} "
                                      "threads"}}


def fp64_single(sk, lps, budget_s: float) -> dict:
    """The plain FP64 engine, one instance per solve, inputs resident, until budget_s."""
    batches = []
    for lp in lps:
        b = sk.Batch(solve_options(sk, "fp64"))
        b.add(lp)
        b.upload()
        batches.append(b)
    batches[0].solve()  # warm (graph capture)
    sk.lib().sk_gpu_synchronize()
    done, iters, t0 = 0, 0, time.perf_counter()
    for b in batches:
        b.solve()
        done += 1
        iters += b.result(0).iterations
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    for b in batches:
        b.close()
    return {"lps_per_s": done / dt, "iters_per_s": iters / dt, "instances": done,
            "seconds": dt}


def serve_stream(sk, requests: int, rate: float, slots: int, max_batch: int,
                 warm: int = 16, lookahead: int = 256, lanes: int = 1) -> dict:
    """Open-loop load: request i arrives at the i-th point of a Poisson process (numpy
    default_rng(42) exponential gaps, mean 1/rate) and is submitted then; the LP (preset 4,
    instance i) is generated ahead by worker threads. Latency = completion - arrival."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor

    gaps = np.random.default_rng(42).exponential(1.0 / rate, requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    srv = sk.Server(solve_options(sk), slots=slots, max_batch=max_batch, keep_solutions=False,
                    lanes=lanes)
    pool = ThreadPoolExecutor(max_workers=3)
    # warm the runtime (module load, every slot's first graph captures and its buffers grown
    # to the large end of the size range) on requests outside the stream: the largest of 256
    # further ones, submitted a few ms apart so that every slot takes some
    side = range(requests, requests + 256)
    big = sorted(side, key=lambda i: -sk.Lp.preset(4, i)["num_rows"])[:warm]
    for i in big:
        srv.submit(sk.Lp.generate(4, i))
        time.sleep(0.003)
    srv.wait()
    for rid in srv.completed():
        srv.result(rid)
    pending, issued = {}, 0

    def refill(upto):
        nonlocal issued
        while issued < min(upto, requests):
            pending[issued] = pool.submit(sk.Lp.generate, 4, issued)
            issued += 1

    refill(lookahead)
    records, results = [], []

    def collect():
        for rid in srv.completed():
            res, rec = srv.result(rid)
            records.append(rec)
            results.append(res)

    t0 = srv.now() + 0.2
    late = 0.0
    for i in range(requests):
        lp = pending.pop(i).result()
        refill(i + 1 + lookahead)
        target = t0 + float(arrivals[i])
        while True:
            d = target - srv.now()
            if d <= 0:
                break
            time.sleep(min(d, 0.002))
        late = max(late, srv.now() - target)
        srv.submit(lp, arrival=target)
        if i % 16 == 0:
            collect()
    srv.wait()
    collect()
    st = srv.stats()
    srv.close()
    pool.shutdown()
    lat = sorted(r["latency"] for r in records)
    first = t0
    last = max(r["finished"] for r in records)
    converged = sum(1 for r in results if r.converged)
    iters = sum(r.iterations for r in results)

    def pct(p):
        return 1e3 * lat[min(len(lat) - 1, max(0, int(round(p * len(lat))) - 1))]

    return {"requests": requests, "rate_per_s": rate, "slots": slots, "lanes": lanes,
            "max_batch": max_batch,
            "completed": len(records), "converged": converged,
            "lps_per_s": converged / (last - first), "pdhg_iters_per_s": iters / (last - first),
            "latency_ms": {"p50": pct(0.50), "p90": pct(0.90), "p99": pct(0.99),
                           "max": 1e3 * lat[-1], "mean": 1e3 * statistics.fmean(lat)},
            "queue_wait_ms_p50": 1e3 * statistics.median(r["admitted"] - r["arrival"] for r in records),
            "queue_wait_ms_p99": 1e3 * sorted(r["admitted"] - r["arrival"] for r in records)[
                min(len(records) - 1, int(0.99 * len(records)))],
            "service_ms": {"p50": 1e3 * statistics.median(r["finished"] - r["admitted"] for r in records),
                           "p99": 1e3 * sorted(r["finished"] - r["admitted"] for r in records)[
                               min(len(records) - 1, int(0.99 * len(records)))]},
            "mean_batch": statistics.fmean(r["batch_size"] for r in records),
            "batches": st["batches"], "max_queue": st["max_queue"],
            "peak_device_bytes": st["peak_held"], "failed": st["failed"],
            "max_submit_lag_ms": 1e3 * late, "seconds": last - first}


def precision_comparison(sk, lp, time_limit: float, use_graph: int) -> dict:
    """The FP32/FP64 split engine against the plain FP64 engine on one LP, at relative KKT
    1e-4 and 1e-6: iterations, solve time and the relative primal, dual and gap errors of
    each answer as the engine's FP64 evaluation on the original model measures them (the
    oracle's evaluate gives the same figures to 1e-7 relative: tests/test_gpu_configs.py)."""
    out = {}
    for tol in (1e-4, 1e-6):
        for prec in ("fp32", "fp64"):
            t0 = time.perf_counter()
            r = sk.solve(lp, solve_options(sk, prec, tolerance=tol, time_limit=time_limit,
                                           use_graph=use_graph))
            out[f"{prec}@{tol:g}"] = {
                "converged": bool(r.converged), "status": r.status_name,
                "iterations": r.iterations, "seconds": time.perf_counter() - t0,
                "rel_primal": r.rel_primal, "rel_dual": r.rel_dual, "rel_gap": r.rel_gap}
    return out


def fp64_finish(sk, lp, first, tolerance: float, use_graph: int) -> dict:
    """The finish of an LP above the interior point's factor cap (sk_pdhg_refine): the FP64
    engine warm-started from the FP32 answer, to `tolerance` relative KKT. Its accuracy
    against a known optimum is tests/test_gpu_finish.py's; here its cost."""
    t0 = time.perf_counter()
    r = sk.refine(lp, first, pdhg_tolerance=tolerance, pdhg_stop_at_request=True,
                  pdhg_absolute_primal=False, use_graph=use_graph)
    return {"tolerance": tolerance, "converged": bool(r.converged), "status": r.status_name,
            "fp64_iterations": r.polish_iterations, "seconds": time.perf_counter() - t0,
            "rel_primal": r.rel_primal, "rel_dual": r.rel_dual, "rel_gap": r.rel_gap,
            "objective_change": abs(r.objective - first.objective) / max(1.0, abs(r.objective)),
            "method": "FP64 PDHG warm-started from the FP32 answer (the dense interior point "
                      "declines an LP of this size)"}


def presolve_cost(sk, lp) -> dict:
    """Host cost of the reference's presolve and postsolve (sk_presolve, sk_postsolve) on the
    workload LP, beside the engine's solve: they run in solve() around every LP engine
    (solve.cpp:423-443). The bench's timed solves run without them, as the reference's own
    PDHG benchmark would be compared engine to engine."""
    import numpy as np
    t0 = time.perf_counter()
    p = sk.presolve(lp)
    t1 = time.perf_counter()
    nr, nc = p.info["reduced_rows"], p.info["reduced_cols"]
    p.postsolve(np.zeros(nc), np.zeros(nr), np.zeros(nc))
    t2 = time.perf_counter()
    return {"presolve_seconds": t1 - t0, "postsolve_seconds": t2 - t1, "passes": p.info["passes"],
            "records": {k: v for k, v in p.info["by_kind"].items() if v},
            "reduced": [nr, nc, p.info["reduced_nonzeros"]]}


def polish_sample(sk, lps, results, count: int) -> dict:
    """The interior-point finish (sk_pdhg_polish) of the first `count` batch answers: the
    step from the 1e-4 first-order answer to an optimal one within 1e-6 relative objective."""
    rows = []
    for lp, first in list(zip(lps, results))[:count]:
        t0 = time.perf_counter()
        r = sk.polish(lp, first)
        rows.append({"seconds": time.perf_counter() - t0, "status": r.status_name,
                     "iterations": r.polish_iterations, "rows": lp.dims[0],
                     "objective_change": abs(r.objective - first.objective) / max(1.0, abs(r.objective))})
    secs = [r["seconds"] for r in rows]
    return {"instances": len(rows), "optimal": sum(r["status"] == "optimal" for r in rows),
            "seconds_mean": statistics.fmean(secs), "seconds_max": max(secs),
            "ipm_iterations_mean": statistics.fmean(r["iterations"] for r in rows),
            "ipm_iterations_max": max(r["iterations"] for r in rows),
            "max_rows": max(r["rows"] for r in rows),
            "first_order_objective_error_max": max(r["objective_change"] for r in rows),
            "method": "PDHG answer at 1e-4 relative KKT, then the interior point on the GPU "
                      "(dense normal equations) to its 1e-8 stop rule"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--fp64-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--spmv-stage", type=int, default=-1,
                    help="staging order: -1 default, 0 CSR order, 1 sorted by gather index")
    ap.add_argument("--tile-nnz", type=int, default=0,
                    help="row-block capacity: 0 the library default (2,560), -1 picked by the "
                         "autotuner at upload, n > 0 that capacity")
    ap.add_argument("--steps-per-block", type=int, default=40,
                    help="PDHG steps per captured graph")
    ap.add_argument("--streams", type=int, default=0,
                    help="sub-batches run side by side: 0 automatic, 1 one block-diagonal batch")
    ap.add_argument("--eval-stride", type=int, default=20,
                    help="evaluation stage after every this many steps of a captured block")
    ap.add_argument("--requests", type=int, default=0,
                    help="serving stream length (0: 10000 for --workload c4, 3000 otherwise)")
    ap.add_argument("--rate", type=float, default=200.0, help="serving arrivals per second")
    ap.add_argument("--slots", type=int, default=2, help="serving: concurrent batches")
    ap.add_argument("--max-batch", type=int, default=256, help="serving: instances per batch")
    ap.add_argument("--lanes", type=int, default=1,
                    help="serving: batches a slot may run at once (the next starts in the tail)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--polish-sample", type=int, default=16,
                    help="batch answers handed to the interior-point finish (0: none)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every kernel from the host instead of replaying captured "
                         "hipGraphs (for profiler runs; same kernels and answers)")
    ap.add_argument("--iteration-limit", type=int, default=-1,
                    help="cap every solve at this many PDHG iterations (counter-profiling runs "
                         "only; capped solves are not the metric)")
    ap.add_argument("--column-slices", type=int, default=0,
                    help="products over XCD slices for a batch of one LP: 0 automatic, 1 on, 2 off")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the FP32/FP64 comparison of the single-LP workloads")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(batches shard by instance across GPUs; that path is not built)", file=sys.stderr)
        return 2

    import sih26_amd as sk
    if not sk.gpu_available():
        print("no HIP device", file=sys.stderr)
        return 1

    if args.workload == "c4":
        return main_serving(sk, args)
    lps = instances(sk, args.workload, args.instances)
    layout = dict(spmv_stage=args.spmv_stage, tile_nnz=args.tile_nnz, eval_stride=args.eval_stride,
                  steps_per_block=args.steps_per_block, use_graph=0 if args.no_graph else 1,
                  iteration_limit=args.iteration_limit, column_slices=args.column_slices)
    batch = sk.Batch(solve_options(sk, streams=args.streams, **layout))
    for lp in lps:
        batch.add(lp)
    batch.upload()
    for _ in range(args.warmup):
        batch.solve()
    sk.lib().sk_gpu_synchronize()

    times = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        batch.solve()
        times.append(time.perf_counter() - t0)
    sk.lib().sk_gpu_synchronize()
    total = time.perf_counter() - t_start

    results = [batch.result(k) for k in range(len(lps))]
    solved = sum(1 for r in results if r.converged)
    iterations = sum(r.iterations for r in results)
    stats = batch.stats()
    n = len(lps)

    # the step kernels as the engine runs them: every sub-batch's launch at once
    split_kernels = {}
    if stats["streams"] > 1:
        for k in (0, 1, 2):
            ms, nbytes = batch.time_kernel(k, 4 * args.kernel_reps)
            split_kernels[KERNEL_NAMES[k]] = {"ms": ms, "GB/s": nbytes / (ms * 1e-3) / 1e9,
                                              "frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    # dominant kernel: time every step kernel with HIP events on the engine's stream, over the
    # whole workload as one block-diagonal batch (one launch covers every instance)
    batch.close()
    # (with the capacity the solve ran with, also when the autotuner picked it)
    whole = sk.Batch(solve_options(sk, streams=1, **dict(layout, tile_nnz=stats["tile_nnz"])))
    for lp in lps:
        whole.add(lp)
    whole.upload()
    whole.solve()
    kern = {}
    for k in (0, 1, 2):
        ms, nbytes = whole.time_kernel(k, args.kernel_reps)
        kern[k] = {"name": KERNEL_NAMES[k], "ms": ms, "bytes": nbytes,
                   "gbs": nbytes / (ms * 1e-3) / 1e9}
        # the same launch after the solve in normal mode: every instance has stopped, so this
        # is the floor cost of a launch whose tiles all exit at once
        kern[k]["stopped_ms"] = whole.time_kernel(k + 3, args.kernel_reps)[0]
    whole.close()
    dom = max((1, 2), key=lambda k: kern[k]["ms"])
    traffic = None
    tf = ROOT / "profiles" / f"traffic_{args.workload}.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get(kern[dom]["name"])

    line = {
        "metric": "LPs solved/sec + PDHG iters/sec at 1e-4 rel KKT; p50/p99 solve latency",
        "value": solved * args.steps / total,
        "unit": "LPs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (SpMV, iterates) + fp64 (KKT, gap, restart reductions)",
        "data": "synthetic (seeded generator in sih26_amd/csrc/generator.cpp)",
        "config": {"workload": WORKLOADS[args.workload], "instances": n,
                   "rows": stats["total_rows"], "cols": stats["total_cols"],
                   "nnz": stats["total_nnz"], "sub_batches": stats["streams"],
                   "tile_nnz": stats["tile_nnz"], "fused_step": bool(stats["fused"]),
                   "column_slices": stats["column_slices"],
                   "tolerance": 1e-4,
                   "stop": "relative KKT (primal, dual, gap) <= 1e-4, FP64 on the original model"},
        "pdhg_iters_per_s": iterations * args.steps / total,
        "solved": solved,
        "instance_iterations": {"sum": iterations, "max": max(r.iterations for r in results),
                                "median": statistics.median(r.iterations for r in results)},
        "batch_solve_ms": {"p50": 1e3 * statistics.median(times),
                           "p99": 1e3 * sorted(times)[max(0, int(round(0.99 * len(times))) - 1)]},
        "phase_ms": {"setup": stats["setup_ms"], "loop": stats["loop_ms"],
                     "report": stats["report_ms"], "graph_launches": stats["graph_launches"]},
        "kernels": {v["name"]: {"ms": v["ms"], "GB/s": v["gbs"], "bytes": v["bytes"],
                                "all_stopped_ms": v["stopped_ms"]}
                    for v in kern.values()},
        "step_kernels_concurrent": {
            "what": f"each kernel launched on all {stats['streams']} sub-batch streams at once, "
                    "wall time per repetition (the schedule the solve runs)",
            **split_kernels} if split_kernels else None,
        "roofline": {"bound": "hbm", "kernel": kern[dom]["name"], "achieved": kern[dom]["gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kern[dom]["gbs"] / HBM_PEAK_GBS, "traffic": traffic},
    }
    if args.polish_sample > 0:
        line["polish"] = polish_sample(sk, lps, results, args.polish_sample)
    if not args.no_fp64 and args.workload == "c3":
        f = fp64_single(sk, lps, args.fp64_budget)
        line["fp64_single"] = f
        line["speedup_vs_fp64_single"] = line["value"] / f["lps_per_s"]
    if args.workload == "c2" and not args.no_compare:
        line["precision_comparison"] = precision_comparison(sk, lps[0], 120.0, layout["use_graph"])
    if args.workload == "c5" and not args.no_compare:
        line["finish"] = fp64_finish(sk, lps[0], results[0], 1e-7, layout["use_graph"])
    if args.workload in ("c2", "c5") and not args.no_compare:
        line["presolve"] = presolve_cost(sk, lps[0])
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(lps, args.cpu_budget)
    if not args.no_serve and args.workload == "c3":
        srv = serve_stream(sk, args.requests or 3000, args.rate, args.slots, args.max_batch,
                           lanes=args.lanes)
        line["serving"] = srv
        line["latency_ms"] = {"p50": srv["latency_ms"]["p50"], "p99": srv["latency_ms"]["p99"],
                              "source": f"the first {srv['requests']} requests of the serving "
                                        "stream (see serving; --workload c4 runs all 10,000)"}
    print(json.dumps(line))
    return 0


def main_serving(sk, args) -> int:
    """--workload c4: the serving stream is the measured workload."""
    srv = serve_stream(sk, args.requests or 10000, args.rate, args.slots, args.max_batch,
                       lanes=args.lanes)
    line = {
        "metric": "LPs solved/sec + PDHG iters/sec at 1e-4 rel KKT; p50/p99 solve latency",
        "value": srv["lps_per_s"],
        "unit": "LPs/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": 1e3 * srv["seconds"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (SpMV, iterates) + fp64 (KKT, gap, restart reductions)",
        "data": "synthetic (seeded generator in sih26_amd/csrc/generator.cpp)",
        "config": {"workload": WORKLOADS["c4"], "requests": srv["requests"],
                   "rate_per_s": args.rate, "slots": args.slots, "max_batch": args.max_batch,
                   "tolerance": 1e-4},
        "pdhg_iters_per_s": srv["pdhg_iters_per_s"],
        "latency_ms": srv["latency_ms"],
        "serving": srv,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_serving(sk, args.rate, args.cpu_budget)
    print(json.dumps(line))
    return 0


def cpu_serving(sk, rate: float, budget_s: float) -> dict:
    """The FP64 CPU oracle as a sequential server on the same request stream (same arrivals,
    same LPs): request i starts when it has arrived and request i-1 is done; its service time
    is the oracle's measured solve time. Run on the first requests until `budget_s` of solving
    is spent, so the latencies are those of that prefix of the stream."""
    import numpy as np
    from oracle import oracle
    oracle.build(ref=False, arch_flags="-march=native")
    oracle.set_threads(1)
    gaps = np.random.default_rng(42).exponential(1.0 / rate, 100000)
    arrivals = np.cumsum(gaps) - gaps[0]
    service, free_at, lat, spent = [], 0.0, [], 0.0
    for i in range(len(arrivals)):
        lp = sk.Lp.generate(4, i).export()
        t0 = time.perf_counter()
        oracle.solve_pdhg(lp, tolerance=1e-4, stop_at_request=True, absolute_primal=False)
        s = time.perf_counter() - t0
        spent += s
        start = max(free_at, float(arrivals[i]))
        free_at = start + s
        service.append(s)
        lat.append(free_at - float(arrivals[i]))
        if spent > budget_s:
            break
    q = lambda v, p: 1e3 * float(np.percentile(v, p))  # noqa: E731
    return {"value": len(service) / spent, "unit": "LPs/s", "cores": 1, "kind": "port",
            "sample": f"the stream's first {len(service)} requests through the FP64 oracle as a "
                      f"sequential server, 1 thread, {spent:.1f} s of solving",
            "service_ms": {"p50": q(service, 50), "p99": q(service, 99)},
            "latency_ms": {"p50": q(lat, 50), "p99": q(lat, 99)},
            "host_cpu": host_cpu()}


if __name__ == "__main__":
    sys.exit(main())
