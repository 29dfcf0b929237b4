#!/usr/bin/env python3
"""FA2 + Barnes-Hut iterations/s on one MI355X: the 1,000,000-node Barabasi-Albert
workload (m=4, positions uniform in [-10,10]^3, theta 0.8).

    python bench.py [--gpus 1] [--steps 200] [--warmup 10] [--no-cpu-baseline]
                    [--no-warm-start] [--copy-probe]
                    [--workload ba_1m|batch_4096|sbm2000|brute_sbm2000|pagerank_ba1m|spmv_ba1m|
                                sgd_rgg200k|traverse|mpnn_ba1m]

A step is one full FA2+BH iteration (octree build, walk, attraction, gravity,
adaptive speed, apply) with inputs already resident in HBM.  The timed region
replays the captured iteration graph; HIP events on the engine stream bracket
every force kernel inside it, which gives the roofline figure.  The CPU
baseline is the oracle's restatement of the reference path timed on this
host on a bounded sample of the same workload.  The second half of the metric
is measured on the same graph and start: time to reach 1.01x the plain run's
converged normalized stress (its stress after 2,000 iterations; 50 max-min
pivot BFS rows), from the learned warm start (prediction wall time included)
versus from the random start, with the plain run as fallback.  Prints one
JSON line.

--copy-probe copies 256 MiB three times with torch and exits: the HBM-counter
calibration run (FETCH_SIZE / WRITE_SIZE of a copy of known size).

The other workloads print one JSON line each for the sibling paths:
  batch_4096     4,096 graphs (50..20,000 nodes) in one batch, 500 iterations,
                 against the same graphs run one engine at a time (sampled);
  sbm2000        FA2+BH on the 2,000-node SBM (theta 1.2, 500 iterations), the
                 reference's CPU-runnable configuration;
  brute_sbm2000  the brute-force engine (fa2-brute-hip) on the 2,000-node SBM, 500
                 iterations;
  pagerank_ba1m  PageRank on the 1M-node BA CSR, d=0.85, L1 tol 1e-6;
  sgd_rgg200k    SGD sparse stress on the 200k-node 2-D RGG, 200 pivots, 30 epochs;
  traverse       single-source BFS on the 200k-node RGG and the 1M-node BA graph,
                 and their connected components;
  mpnn_ba1m      the warm-start network: inference at 1M nodes, a training step at 100k.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "jump-cannon_amd")]

import numpy as np  # noqa: E402

METRIC = "FA2+BH iters/s at 1M nodes; time-to-stress≤1.01× plain: learned vs plain"
HBM_PEAK_GBS = 8000.0
VALU_FP32_PEAK_TFLOPS = 157.3
PMC_SUMMARY = os.path.join(ROOT, "profiles", "fa2_force_ba1m_pmc.json")


def pmc_traffic(name: str):
    """Calibrated HBM bytes per launch of a workload's dominant kernel from its
    committed counter summary (profiles/<name>_pmc.json), or None."""
    path = os.path.join(ROOT, "profiles", f"{name}_pmc.json")
    try:
        with open(path) as fh:
            return int(json.load(fh)["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def _threads() -> int:
    return min(16, os.cpu_count() or 1)


def _timed_rate(fn, budget_s: float):
    """Calls fn(k) -> units done, growing k until about budget_s of work;
    returns (units/s, units)."""
    t0 = time.perf_counter()
    done = fn(1)
    first = time.perf_counter() - t0
    more = max(0, int(budget_s / max(first, 1e-4)) - 1)
    total = first
    if more:
        t1 = time.perf_counter()
        done += fn(more)
        total += time.perf_counter() - t1
    return done / total, done


def log(msg: str) -> None:
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def force_kernel_bytes(n: int, n_tree: int, nnz_att: int) -> int:
    """Algorithmic HBM bytes of one force launch: every input read once, every
    output written once.  Per body: position 12, F_{t-1} 12, F write 12,
    degree 4, Morton permutation 4; per octree node 24 (com+mass 16, info+rope 8)
    read once; the attraction lists: offsets 4(n+1), neighbour ids 4 per entry,
    neighbour positions 12 per entry."""
    return 44 * n + 24 * n_tree + 4 * (n + 1) + 16 * nnz_att


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_rate(off, nb, xyz, s, threads: int, budget_s: float):
    """Iterations/s of the oracle's reference-path restatement from iteration 0."""
    from oracle import oracle as O

    t0 = time.perf_counter()
    x, old, se, tm = O.fa2_run(off, nb, xyz, 1, s, tree_kind=0, exact=False, threads=threads)
    first = time.perf_counter() - t0
    more = max(0, int(budget_s / max(first, 1e-3)) - 1)
    iters, total = 1, first
    if more:
        t1 = time.perf_counter()
        O.fa2_run(off, nb, x, more, s, tree_kind=0, exact=False, threads=threads,
                  old_force=old, se=se)
        iters += more
        total = first + (time.perf_counter() - t1)
    return iters / total, iters


def cpu_baseline(off, nb, xyz, theta: float, budget_s: float) -> dict:
    from oracle import oracle as O

    threads = min(16, os.cpu_count() or 1)
    s = O.Fa2Settings.make(theta=theta)
    rate, iters = _cpu_rate(off, nb, xyz, s, threads, budget_s)
    rate1, iters1 = _cpu_rate(off, nb, xyz, s, 1, budget_s / 2)
    return {"value": rate, "unit": "iterations/s", "cores": threads, "kind": "port",
            "single_thread_value": rate1, "cpu_model": _cpu_model(),
            "nproc": os.cpu_count(),
            "sample": f"{iters} iterations ({threads} threads) and {iters1} iterations "
                      f"(1 thread) of the same 1M-node workload from iteration 0: "
                      f"reference insertion octree built serially, walk/attraction/apply "
                      f"under OpenMP, f64 node-order speed sums"}


def copy_probe() -> int:
    import torch

    a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copied 256 MiB three times")
    return 0


def warm_start(off, nb, xyz, theta: float) -> dict | None:
    from jumpcannon import warmstart as W

    if not os.path.exists(W.WEIGHTS):
        log("no trained warm-start weights; skipping the learned half of the metric")
        return None
    net, coef = W.load()
    r = W.evaluate(net, coef, off, nb, xyz, theta=theta, log=log)
    keep = ("horizon", "check_every", "plain_init_stress", "plain_stress_final", "target_stress",
            "plain_iters_to_target", "plain_ms_to_target", "learned_init_stress",
            "learned_stress_final", "learned_iters_to_target", "inference_ms",
            "learned_ms_to_target", "fallback_to_plain", "speedup",
            "scaled_random_iters_to_target", "scaled_random_ms_to_target", "multilevel_levels",
            "multilevel_descent_ms", "multilevel_level0_start_stress", "multilevel_level0_iters_to_target",
            "multilevel_ms_to_target", "plain_init_cv", "plain_cv_final", "target_cv", "learned_init_cv",
            "learned_cv_final", "plain_iters_to_both", "plain_ms_to_both", "plain_stays_within_both",
            "learned_iters_to_both", "learned_ms_to_both", "learned_stays_within_both",
            "learned_max_stress_after_target", "speedup_both", "scaled_random_iters_to_both",
            "scaled_random_ms_to_both", "multilevel_level0_iters_to_both", "multilevel_ms_to_both")
    return {k: (round(r[k], 5) if isinstance(r[k], float) else r[k]) for k in keep}


def _line(metric, value, unit, steps, warmup, ms_per_step, dtype, config, **extra) -> int:
    out = {"metric": metric, "value": round(value, 4), "unit": unit, "n_gpus": 1,
           "steps": steps, "warmup": warmup, "ms_per_step": round(ms_per_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic", "config": config}
    out.update(extra)
    print(json.dumps(out))
    return 0


def bench_batch(args) -> int:
    """4,096 graphs, 500 iterations in one batch; the same graphs one engine at a
    time, timed on every 16th graph and scaled by 16."""
    import jumpcannon as jc
    from jumpcannon import workloads

    iters = args.batch_iters
    t = time.perf_counter()
    graphs, pos = workloads.batch_4096()
    total = sum(len(o) - 1 for o, _ in graphs)
    log(f"workload: {len(graphs)} graphs, {total} nodes ({time.perf_counter() - t:.1f}s)")
    b = jc.Fa2BhBatch(theta=1.0)
    b.init(graphs, pos)
    b.step(2, readback=False)
    sched = b.schedule()
    b_ms, bs = b.step_timed(iters)
    log(f"batch: {iters} iterations in {b_ms:.1f} ms (schedule {sched[0]}, force block {bs}, "
        f"centre of mass from depth {sched[3]} in chunks of {sched[4]}, sort key {sched[5]} bits)")
    seq_ms = 0.0
    stride = max(1, args.seq_stride)
    sample = range(0, len(graphs), stride) if not args.no_sequential else range(0)
    t_seq = time.perf_counter()
    for g in sample:
        off, nb = graphs[g]
        e = jc.Fa2BhEngine(theta=1.0)
        e.init(off, nb, pos[g])
        e.step(2, readback=False)
        _, it_ms = e.step_timed(20)
        seq_ms += it_ms / 20 * iters
        e.close()
    seq_ms *= len(graphs) / max(len(sample), 1)
    which = f"all {len(graphs)} graphs" if stride == 1 else f"every {stride}-th graph ({len(sample)} of {len(graphs)})"
    log(f"one engine per graph ({which}, {time.perf_counter() - t_seq:.0f}s): {seq_ms:.1f} ms for {iters} iterations")
    return _line("batched FA2+BH node-iterations/s (4,096 graphs)", total * iters / (b_ms / 1e3),
                 "node-iterations/s", iters, 2, b_ms / iters, "f32",
                 {"workload": "batch_4096_fa2bh", "graphs": len(graphs), "total_nodes": total,
                  "sizes": "log-uniform 50..20000, half SBM half BA", "theta": 1.0,
                  "parallelism": "single-gpu"},
                 batch_ms=round(b_ms, 2), force_block=bs, schedule=sched[0],
                 tuned={"force_block": sched[2], "com_fuse_depth": sched[3], "com_chunk": sched[4],
                        "sort_key_bits": sched[5]},
                 sequential_ms_estimate=round(seq_ms, 1) if len(sample) else None,
                 speedup_vs_sequential=round(seq_ms / b_ms, 2) if len(sample) else None,
                 sequential_sample=(f"{which}, each its own engine, 20 timed iterations after 2, scaled to "
                                    f"{iters} iterations" + ("" if stride == 1 else f" and x{stride}")))


def bench_sbm2000(args) -> int:
    """FA2+BH on the 2,000-node SBM (theta 1.2, 500 iterations): the reference's own
    CPU-runnable configuration, against the oracle's restatement of its CPU path
    (insertion octree, serial, f64 node-order sums) on one host thread."""
    import jumpcannon as jc
    from jumpcannon import workloads

    iters = 500
    off, nb, xyz = workloads.sbm_2000()
    n = len(off) - 1
    e = jc.Fa2BhEngine(theta=1.2)
    e.init(off, nb, xyz)
    e.step(5, readback=False)
    e.sync()
    t = time.perf_counter()
    e.step(iters, readback=False)
    e.sync()
    wall = time.perf_counter() - t
    f_ms, it_ms = e.step_timed(iters)
    ph = e.profile(args.profile_iters)
    log(f"sbm2000: {iters} iterations in {wall * 1e3:.1f} ms (graph replay), walk mode {e.walk_mode()}; "
        f"timed: force {f_ms / iters * 1e3:.1f} us, iteration {it_ms / iters * 1e3:.1f} us")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        def run(k):
            O.fa2_run(off, nb, xyz, k, O.Fa2Settings.make(theta=1.2), tree_kind=0, exact=False, threads=1)
            return k
        rate, done = _timed_rate(run, args.cpu_seconds / 2)
        cpu = {"value": rate, "unit": "iterations/s", "cores": 1, "kind": "port", "cpu_model": _cpu_model(),
               "nproc": os.cpu_count(),
               "sample": f"{done} iterations of the oracle's restatement of the reference CPU path "
                         f"(insertion octree fa2_bh.rs:747-1045 built serially, walk, speed, apply) on one thread"}
        log(f"sbm2000 cpu baseline {rate:.1f} iterations/s on 1 thread")
    return _line("FA2+BH iterations/s at 2,000 nodes (SBM, theta 1.2)", iters / wall, "iterations/s", iters, 5,
                 wall * 1e3 / iters, "f32",
                 {"workload": "sbm2000_fa2bh", "n_nodes": n, "csr_nnz": len(nb), "theta": 1.2,
                  "parallelism": "single-gpu"},
                 phase_ms={k: round(v, 4) for k, v in ph.items()}, walk_mode=e.walk_mode(),
                 cpu_baseline=cpu)


def bench_brute(args) -> int:
    """fa2-brute-hip (fa2_brute.rs) on the 2,000-node SBM: 500 iterations of the
    O(n^2) engine, iterations/s, with the repulsion kernel's flop rate."""
    import jumpcannon as jc
    from jumpcannon import workloads

    iters = 500
    off, nb, xyz = workloads.sbm_2000()
    n = len(off) - 1
    e = jc.Fa2BruteEngine()
    e.init(off, nb, xyz)
    e.step(5, readback=False)
    e.sync()
    mode = e.walk_mode()
    t = time.perf_counter()
    e.step(iters, readback=False)  # the captured 2-iteration graph, as a caller's step(k) runs
    e.sync()
    wall = time.perf_counter() - t
    f_ms, it_ms = e.step_timed(iters)  # events around every force kernel
    # per pair (force_atlas2.wgsl:82-91): d (3), |d|^2 (5), max (1), coefficient (3), F += d*c (6)
    flops = 18.0 * n * (n - 1)
    achieved = flops / (f_ms / iters / 1e3) / 1e12
    log(f"brute: {iters} iterations in {wall * 1e3:.1f} ms (graph replay); with events around each force "
        f"kernel {it_ms:.1f} ms, force kernels {f_ms:.1f} ms; repulsion kernel "
        f"{['one lane', '4 lanes', '16 lanes'][mode]} per body")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        def run(k):
            O.fa2_run(off, nb, xyz, k, O.Fa2Settings.make(), tree_kind=2, exact=False, threads=1)
            return k
        rate, done = _timed_rate(run, args.cpu_seconds / 2)
        cpu = {"value": rate, "unit": "iterations/s", "cores": 1, "kind": "port", "cpu_model": _cpu_model(),
               "nproc": os.cpu_count(),
               "sample": f"{done} iterations of the oracle's restatement of the brute-force iteration "
                         f"(force_atlas2.wgsl:69-165, fa2_speed.rs) on one thread"}
        log(f"brute cpu baseline {rate:.1f} iterations/s on 1 thread")
    return _line("FA2 brute-force iterations/s (2,000-node SBM)", iters / wall, "iterations/s",
                 iters, 5, wall * 1e3 / iters, "f32",
                 {"workload": "brute_sbm2000", "n_nodes": n, "csr_nnz": len(nb), "engine": "fa2-brute-hip",
                  "parallelism": "single-gpu"},
                 roofline={"bound": "valu", "achieved": round(achieved, 3), "peak": VALU_FP32_PEAK_TFLOPS,
                           "unit": "TFLOP/s", "frac": round(achieved / VALU_FP32_PEAK_TFLOPS, 5),
                           "traffic": None, "flops_per_launch": flops,
                           "avg_launch_ms": round(f_ms / iters, 5),
                           "kernel": ["k_force_brute<256>", "k_force_brute_grp<256, 4>",
                                      "k_force_brute_grp<256, 16>"][mode]},
                 cpu_baseline=cpu)


def bench_pagerank(args) -> int:
    """PageRank on the 1M-node BA CSR to L1 < 1e-6; iterations/s of the fixed
    per-iteration kernels and the HBM figure of one iteration."""
    import jumpcannon as jc
    from jumpcannon import workloads

    off, nb, _ = workloads.ba_1m(args.nodes)
    n, nnz = len(off) - 1, len(nb)
    plan = jc.PageRankPlan(off, nb, 0.85)
    plan.run(100, 1e-6, readback=False)
    _, iters, resid, run_ms = plan.run(100, 1e-6, readback=False)
    ms_it = plan.time_iterations(100)
    # column ids and the gathered x = rank/deg per entry, offsets, the rank
    # write, and the x write plus its compulsory read: 8 nnz + 16 n
    algo = 8 * nnz + 16 * n
    achieved = algo / (ms_it / 1e3) / 1e9
    log(f"pagerank: {iters} iterations to L1 {resid:.3e} in {run_ms:.3f} ms; {ms_it:.4f} ms/iteration")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        def it_rate(threads):
            return _timed_rate(lambda k: O.pagerank_pull(off, nb, 0.85, k, -1.0, threads=threads)[1],
                               args.cpu_seconds / 2)
        rate, done = it_rate(_threads())
        rate1, done1 = it_rate(1)
        cpu = {"value": rate, "unit": "iterations/s", "cores": _threads(), "kind": "port",
               "single_thread_value": rate1, "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
               "sample": f"{done} pull iterations ({_threads()} threads) and {done1} (1 thread) of the "
                         f"oracle's restatement of cpu/gpu_pagerank's update (pagerank.rs:298-331, "
                         f"pagerank.wgsl:30-46) on the same CSR, exact f64 residual, stopping rule off"}
        log(f"pagerank cpu baseline {rate:.1f} it/s on {_threads()} threads, {rate1:.1f} on 1")
    return _line("PageRank iterations/s at 1M nodes (BA m=4)", 1e3 / ms_it, "iterations/s", 100, 100,
                 ms_it, "f32",
                 {"workload": "pagerank_ba1m", "n_nodes": n, "csr_nnz": nnz, "damping": 0.85,
                  "tol_l1": 1e-6, "parallelism": "single-gpu"},
                 iterations_to_tol=iters, residual_l1=resid, ms_to_tol=round(run_ms, 4),
                 roofline={"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                           "traffic": pmc_traffic("pagerank_ba1m"), "bytes_per_iteration": algo,
                           "kernel": "k_pr_iter + k_pr_reduce"},
                 cpu_baseline=cpu)


def bench_spmv(args) -> int:
    """y = A x on the 1M-node BA CSR with PageRank's per-entry weights 1/deg(u)
    (the SpMV form of one pull step): products/s of the serial f32 product, with
    the f16-storage and hybrid variants timed beside it."""
    import jumpcannon as jc
    from jumpcannon import workloads

    off, nb, _ = workloads.ba_1m(args.nodes)
    n, nnz = len(off) - 1, len(nb)
    deg = np.diff(off).astype(np.float32)
    w = (np.float32(1.0) / deg)[nb].astype(np.float32)
    x = np.random.default_rng(1).uniform(0, 2.0 / n, n).astype(np.float32)
    times = {}
    for name, kind in (("f32", jc.SPMV_F32), ("f16", jc.SPMV_F16), ("hybrid", jc.SPMV_HYBRID)):
        plan = jc.SpmvPlan(off, nb, w, kind)
        plan.run(x)
        times[name] = plan.time_products(100)
        del plan
    ms = times["f32"]
    # offsets, column ids, f32 weights, x and y once each
    algo = 4 * (n + 1) + 8 * nnz + 8 * n
    achieved = algo / (ms / 1e3) / 1e9
    log("spmv: " + ", ".join(f"{k} {v:.4f} ms" for k, v in times.items()))
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        def products(k):
            for _ in range(k):
                O.spmv(off, nb, w, x)
            return k
        rate, done = _timed_rate(products, args.cpu_seconds / 2)
        cpu = {"value": rate, "unit": "products/s", "cores": 1, "kind": "port", "cpu_model": _cpu_model(),
               "nproc": os.cpu_count(),
               "sample": f"{done} products of the oracle's restatement of cpu_spmv (spmv.rs:88-99): rows "
                         f"in order, each summed in CSR order in f32, one thread as in the reference"}
        log(f"spmv cpu baseline {rate:.1f} products/s on 1 thread")
    return _line("SpMV products/s at 1M nodes (BA m=4, PageRank weights)", 1e3 / ms, "products/s", 100,
                 1, ms, "f32",
                 {"workload": "spmv_ba1m", "n_nodes": n, "csr_nnz": nnz, "parallelism": "single-gpu"},
                 ms_per_product={k: round(v, 5) for k, v in times.items()},
                 roofline={"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                           "traffic": pmc_traffic("spmv_ba1m"), "bytes_per_product": algo,
                           "kernel": "k_spmv_rows<float>"},
                 cpu_baseline=cpu)


def bench_sgd(args) -> int:
    """SGD sparse stress on the 200k-node RGG: 200 pivots, 30 epochs."""
    import jumpcannon as jc
    from jumpcannon import workloads

    off, nb, _ = workloads.rgg_200k()
    n = len(off) - 1
    xyz = workloads.plane_init(n)
    e = jc.SgdStressEngine(n_pivots=200, eta_max=0.0, eta_min=0.01, n_anneal_steps=30)
    t = time.perf_counter()
    e.init(off, nb, xyz)
    init_s = time.perf_counter() - t
    warm = 2  # the first launches load the kernel; epochs cost the same at any eta
    e.time_steps(warm)
    ms = e.time_steps(30)
    k = 200
    algo = 2 * n * k + 32 * n + 16 * k
    achieved = algo / (ms / 30 / 1e3) / 1e9
    log(f"sgd: 30 epochs in {ms:.3f} ms (init incl. BFS {init_s:.2f} s)")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        piv, dist, eta_max = e.pivots()

        def ep_rate(threads):
            def run(kk):
                p = xyz
                for t in range(kk):
                    p = O.sgd_sweep(piv, dist, p, O.anneal_eta(t, eta_max, 0.01, 30), threads=threads)
                return kk
            return _timed_rate(run, args.cpu_seconds / 2)
        rate, done = ep_rate(_threads())
        rate1, done1 = ep_rate(1)
        cpu = {"value": rate, "unit": "epochs/s", "cores": _threads(), "kind": "port",
               "single_thread_value": rate1, "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
               "sample": f"{done} epochs ({_threads()} threads) and {done1} (1 thread) of the oracle's "
                         f"Jacobi sweep (sgd_stress.wgsl:34-62) over the same 200 pivot rows, from the "
                         f"same start"}
        log(f"sgd cpu baseline {rate:.2f} epochs/s on {_threads()} threads, {rate1:.2f} on 1")
    return _line("SGD sparse-stress epochs/s at 200k nodes (RGG, 200 pivots)", 30 / (ms / 1e3),
                 "epochs/s", 30, warm, ms / 30, "f32",
                 {"workload": "sgd_rgg200k", "n_nodes": n, "csr_nnz": len(nb), "pivots": k,
                  "parallelism": "single-gpu"},
                 init_seconds_incl_bfs=round(init_s, 3),
                 roofline={"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                           "traffic": pmc_traffic("sgd_rgg200k"), "bytes_per_epoch": algo,
                           "kernel": "k_sgd_sweep"},
                 cpu_baseline=cpu)


def bench_mpnn(args) -> int:
    """The warm-start network on the 1M-node BA graph: one inference pass
    (forward + on-device centring and scaling, as the warm start runs it) with
    the committed weights, and one training step (forward, loss, backward, Adam)
    on a 100k-node BA graph against its plain FA2 layout.  Device times from HIP
    events for inference (host clock for the training step); the roofline is
    the pass's algorithmic bytes: the CSR and one gathered bf16 feature row per
    entry in each layer's aggregation, the GEMMs' activations in and out, the
    head and the centring."""
    import jumpcannon as jc
    from jumpcannon import warmstart as W, workloads

    off, nb, xyz = workloads.ba_1m(args.nodes)
    n, nnz = len(off) - 1, len(nb)
    net, coef = W.load()
    t = time.perf_counter()
    net.set_graph(off, nb, feat_seed=7, walks=16, noise=np.asarray(xyz) / W.INIT_HALF)
    feat_s = time.perf_counter() - t
    net.forward_scaled(W.size_of(coef, n, nnz))  # first launch loads the kernels
    fwd = [net.forward_scaled(W.size_of(coef, n, nnz))[1] for _ in range(5)]
    fwd_ms = float(np.median(fwd))
    log(f"mpnn: features {feat_s:.2f} s, inference {fwd_ms:.3f} ms at {n} nodes")
    # one training step on 100k nodes
    toff, tnb = jc.ba_graph(100_000, 4, 0xC0FFEE + 7)
    tn = len(toff) - 1
    tx0 = jc.uniform_cube(tn, W.INIT_HALF, 11)
    tgt = jc.Fa2BhEngine(theta=1.0)
    tgt.init(toff, tnb, tx0)
    target = tgt.step(300)
    tnet = jc.Mpnn(seed=3)
    tnet.set_graph(toff, tnb, feat_seed=7, walks=16, noise=tx0 / W.INIT_HALF)
    tnet.set_target(target, n_anchor=32, seed=5)
    tnet.train(2, 1e-3)
    t = time.perf_counter()
    steps = 20
    tnet.train(steps, 1e-3)
    train_ms = (time.perf_counter() - t) * 1e3 / steps
    log(f"mpnn: training step {train_ms:.3f} ms at {tn} nodes (host clock, includes the loss read-back)")
    # per layer: the aggregation reads the CSR and gathers one bf16 feature row
    # per entry (16 features in layer 0, 128 after), writes one mean row per
    # node; the GEMM reads [h | mean] and writes h'; the head reads h and writes
    # 3 f32 per node; centring reads and writes the 3 coordinates
    csr = 4 * (n + 1) + 4 * nnz
    agg_bytes = (csr + 32 * nnz + 32 * n) + 2 * (csr + 256 * nnz + 256 * n)
    gemm_bytes = (64 * n + 256 * n) + 2 * (512 * n + 256 * n)
    pass_bytes = agg_bytes + gemm_bytes + 256 * n + 12 * n + 3 * 12 * n
    achieved = pass_bytes / (fwd_ms / 1e3) / 1e9
    return _line("MPNN warm-start inference passes/s at 1M nodes (BA m=4)", 1e3 / fwd_ms, "passes/s", 5, 1,
                 fwd_ms, "bf16",
                 {"workload": "mpnn_ba1m", "n_nodes": n, "csr_nnz": nnz, "layers": 3, "hidden": 128,
                  "train_workload": "ba_100k_m4", "parallelism": "single-gpu"},
                 features_seconds=round(feat_s, 3), train_step_ms=round(train_ms, 3),
                 roofline={"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                           "bytes_per_pass": pass_bytes, "aggregation_bytes": agg_bytes, "gemm_bytes": gemm_bytes,
                           "note": "algorithmic bytes of the whole inference pass over its device time"})


def bench_traverse(args) -> int:
    """Single-source BFS on the 200k-node RGG (the stress workload's pivot rows:
    hundreds of thin levels, one workgroup) and on the 1M-node BA graph (a few
    levels of hundreds of thousands of nodes, chip-wide), plus connected
    components.  Device time per search from HIP events, median over sources."""
    import jumpcannon as jc
    from jumpcannon import workloads

    out = {}
    for name, (off, nb) in (("rgg_200k", workloads.rgg_200k()[:2]), ("ba_1m", workloads.ba_1m(args.nodes)[:2])):
        n, nnz = len(off) - 1, len(nb)
        plan = jc.BfsPlan(off, nb)
        d, _ = plan.run(0)
        levels = int(d[d != 0xFFFFFFFF].max()) + 1
        times = []
        for src in range(0, n, max(n // 16, 1)):
            times.append(plan.run(src, readback=False)[1])
        ms = float(np.median(times))
        t = time.perf_counter()
        jc.connected_components(off, nb)
        cc_ms = (time.perf_counter() - t) * 1e3
        out[name] = {"n_nodes": n, "csr_nnz": nnz, "levels_from_0": levels, "searches": len(times),
                     "bfs_ms": round(ms, 4),
                     "teps": round((nnz / 2) / (ms / 1e3), 1), "us_per_level": round(1e3 * ms / levels, 3),
                     "components_wall_ms_incl_upload": round(cc_ms, 2)}
        log(f"bfs {name}: {ms:.3f} ms per search ({levels} levels from node 0), "
            f"{(nnz / 2) / (ms / 1e3) / 1e9:.3f} G edges/s; components {cc_ms:.1f} ms")
    r = out["rgg_200k"]
    algo = 132 * r["n_nodes"]  # one padded 128-byte row per node, one 4-byte distance
    achieved = algo / (r["bfs_ms"] / 1e3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        off, nb = workloads.rgg_200k()[:2]
        srcs = list(range(0, r["n_nodes"], max(r["n_nodes"] // 16, 1)))

        def searches(k):
            for q in range(k):
                O.bfs(off, nb, srcs[q % len(srcs)])
            return k
        rate, done = _timed_rate(searches, args.cpu_seconds / 2)
        cpu = {"value": rate * (r["csr_nnz"] / 2), "unit": "edges/s", "cores": 1, "kind": "port",
               "searches_per_s": rate, "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
               "sample": f"{done} single-source searches of the oracle's restatement of cpu_bfs "
                         f"(bfs.rs:44-66, a FIFO queue over the out-rows) on the same RGG and sources, "
                         f"one thread as in the reference"}
        log(f"bfs cpu baseline {rate:.1f} searches/s on 1 thread")
    return _line("BFS traversed edges/s (200k-node RGG, one search)", r["teps"], "edges/s", r["searches"],
                 1, r["bfs_ms"], "u32",
                 {"workload": "bfs_rgg200k", "n_nodes": r["n_nodes"], "csr_nnz": r["csr_nnz"],
                  "parallelism": "single-gpu"},
                 graphs=out,
                 roofline={"bound": "latency", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None,
                           "bytes_per_search": algo,
                           "limiter": "level-synchronous: one dependent row load and one barrier per level "
                                      "in one workgroup"},
                 cpu_baseline=cpu)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nodes", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--profile-iters", type=int, default=4)
    ap.add_argument("--no-warm-start", action="store_true")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--batch-iters", type=int, default=500)
    ap.add_argument("--seq-stride", type=int, default=1,
                    help="batch_4096: time one engine per graph on every k-th graph (1: all 4,096)")
    ap.add_argument("--no-sequential", action="store_true",
                    help="batch_4096: skip the one-engine-per-graph estimate")
    ap.add_argument("--workload", default="ba_1m",
                    choices=["ba_1m", "batch_4096", "brute_sbm2000", "sbm2000", "pagerank_ba1m",
                             "spmv_ba1m", "sgd_rgg200k", "traverse", "mpnn_ba1m"])
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the layout path is single-GPU in this build).", file=sys.stderr)
        return 2
    if args.workload != "ba_1m":
        return {"batch_4096": bench_batch, "brute_sbm2000": bench_brute, "sbm2000": bench_sbm2000,
                "pagerank_ba1m": bench_pagerank,
                "spmv_ba1m": bench_spmv,
                "sgd_rgg200k": bench_sgd, "traverse": bench_traverse,
                "mpnn_ba1m": bench_mpnn}[args.workload](args)

    import jumpcannon as jc
    from jumpcannon import workloads

    theta = 0.8
    t = time.perf_counter()
    off, nb, xyz = workloads.ba_1m(args.nodes)
    n = len(off) - 1
    log(f"workload: BA n={n} nnz={len(nb)} ({time.perf_counter() - t:.1f}s)")
    e = jc.Fa2BhEngine(theta=theta)
    e.init(off, nb, xyz)
    # attraction entries = 2 x undirected edges (v<u pairs)
    nnz_att = int(np.count_nonzero(np.repeat(np.arange(n, dtype=np.uint32), np.diff(off)) < nb)) * 2

    e.step(args.warmup, readback=False)
    e.sync()
    walk = e.walk_mode(-1)   # re-time the walk variants on the warmed-up layout
    log(f"warmup {args.warmup} done; walk mode {walk}")

    t0 = time.perf_counter()
    force_ms, iter_ms = e.step_timed(args.steps)   # synchronizes at the end
    wall = time.perf_counter() - t0
    value = args.steps / wall
    log(f"timed {args.steps} iterations: {wall * 1e3:.1f} ms, force kernels {force_ms:.1f} ms")

    # the profiled iterations also count the walk's interactions (a separate
    # counting walk after them; plain iterations do not count)
    phases = e.profile(max(args.profile_iters, 1))
    log("phases (ms): " + ", ".join(f"{k}={v:.3f}" for k, v in phases.items()))
    st = e.last_stats()

    fb = force_kernel_bytes(n, st["n_tree"], nnz_att)
    f_avg_s = force_ms / args.steps / 1e3
    achieved = fb / f_avg_s / 1e9
    traffic = None
    if os.path.exists(PMC_SUMMARY):
        with open(PMC_SUMMARY) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    # secondary bound: FP32 vector work of the walk, about 20 FLOP per interaction
    # (difference, squared distance, quotient, accumulation), against the vector
    # FP32 peak; the counters' VALU busy share is in the PMC profile
    inter_flops = 20.0 * st["interactions"]
    tflops = inter_flops / f_avg_s / 1e12
    roofline_valu = {"bound": "valu-fp32", "achieved": round(tflops, 3),
                     "peak": VALU_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tflops / VALU_FP32_PEAK_TFLOPS, 5),
                     "flop_per_interaction": 20, "valu_busy_pmc": 0.54}
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": "k_force (+ k_heavy_terms, k_force_heavy)", "bytes_per_launch": fb,
                "avg_launch_ms": round(force_ms / args.steps, 4),
                "limiter": "vector-memory path (divergent 16-byte node gathers): TA 81% / TD 87% "
                           "busy, VALU 54% (profiles/fa2_force_ba1m_pmc.md)"}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(off, nb, xyz, theta, args.cpu_seconds)
        log(f"cpu baseline {cpu['value']:.3f} it/s on {cpu['cores']} threads")

    ws = None if args.no_warm_start else warm_start(off, nb, xyz, theta)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "ba_1m_fa2bh", "n_nodes": n, "csr_nnz": int(len(nb)),
                   "m": 4, "positions": "uniform [-10,10]^3", "theta": theta,
                   "parallelism": "single-gpu"},
        "interactions_per_iter": st["interactions"],
        "interactions_per_s": round(st["interactions"] * value, 1),
        "octree_nodes": st["n_tree"],
        "walk_mode": walk,
        "phase_ms": {k: round(v, 4) for k, v in phases.items()},
        "roofline": roofline,
        "roofline_secondary": roofline_valu,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(value / cpu["value"], 2) if cpu else None,
        "warm_start": ws,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
