#!/usr/bin/env python3
"""FA2 + Barnes-Hut iterations/s on one MI355X: the 1,000,000-node Barabasi-Albert
workload (m=4, positions uniform in [-10,10]^3, theta 0.8).

    python bench.py [--gpus 1] [--steps 200] [--warmup 10] [--no-cpu-baseline]

A step is one full FA2+BH iteration (octree build, walk, attraction, gravity,
adaptive speed, apply) with inputs already resident in HBM.  The timed region
replays the captured iteration graph; HIP events on the engine stream bracket
every force kernel inside it, which gives the roofline figure.  The CPU
baseline is the oracle's restatement of the reference path timed on this
host on a bounded sample of the same workload.  Prints one JSON line.
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
PMC_SUMMARY = os.path.join(ROOT, "profiles", "fa2_force_ba1m_pmc.json")


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


def cpu_baseline(off, nb, xyz, theta: float, budget_s: float) -> dict:
    from oracle import oracle as O

    threads = min(16, os.cpu_count() or 1)
    s = O.Fa2Settings.make(theta=theta)
    t0 = time.perf_counter()
    x, old, se, tm = O.fa2_run(off, nb, xyz, 1, s, tree_kind=0, exact=False, threads=threads)
    first = time.perf_counter() - t0
    more = max(0, int(budget_s / max(first, 1e-3)) - 1)
    iters = 1
    if more:
        t1 = time.perf_counter()
        O.fa2_run(off, nb, x, more, s, tree_kind=0, exact=False, threads=threads,
                  old_force=old, se=se)
        iters += more
        total = first + (time.perf_counter() - t1)
    else:
        total = first
    return {"value": iters / total, "unit": "iterations/s", "cores": threads, "kind": "port",
            "sample": f"{iters} iterations of the same 1M-node workload from iteration 0: "
                      f"reference insertion octree built serially, walk/attraction/apply "
                      f"under OpenMP ({threads} threads), f64 node-order speed sums"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nodes", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--profile-iters", type=int, default=4)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the layout path is single-GPU in this build).", file=sys.stderr)
        return 2

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
    log(f"warmup {args.warmup} done")

    t0 = time.perf_counter()
    force_ms, iter_ms = e.step_timed(args.steps)   # synchronizes at the end
    wall = time.perf_counter() - t0
    value = args.steps / wall
    st = e.last_stats()
    log(f"timed {args.steps} iterations: {wall * 1e3:.1f} ms, force kernels {force_ms:.1f} ms")

    phases = e.profile(args.profile_iters) if args.profile_iters else {}
    if phases:
        log("phases (ms): " + ", ".join(f"{k}={v:.3f}" for k, v in phases.items()))

    fb = force_kernel_bytes(n, st["n_tree"], nnz_att)
    f_avg_s = force_ms / args.steps / 1e3
    achieved = fb / f_avg_s / 1e9
    traffic = None
    if os.path.exists(PMC_SUMMARY):
        with open(PMC_SUMMARY) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": "k_force", "bytes_per_launch": fb,
                "avg_launch_ms": round(force_ms / args.steps, 4)}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(off, nb, xyz, theta, args.cpu_seconds)
        log(f"cpu baseline {cpu['value']:.3f} it/s on {cpu['cores']} threads")

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
        "phase_ms": {k: round(v, 4) for k, v in phases.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(value / cpu["value"], 2) if cpu else None,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
