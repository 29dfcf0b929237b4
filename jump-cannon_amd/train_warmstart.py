#!/usr/bin/env python3
"""Trains the FA2 warm-start network on the GPU and writes its weights.

    python jump-cannon_amd/train_warmstart.py [--graphs 256] [--epochs 6] [--out PATH]

Training set: seeded SBM/BA graphs of 1k-20k nodes, each with a random initial
layout x0 in [-10,10]^3.  Shape targets are the layouts the SGD sparse-stress
engine reaches from x0; the size law comes from the plain FA2+BH layouts after
--fa2-iters iterations (all graphs in one batched run).  Every forward,
backward and Adam step runs in the library's own HIP kernels.  The default
output is the file the warm start loads (jumpcannon/warmstart_weights.npz).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import jumpcannon as jc  # noqa: E402
from jumpcannon import warmstart as W  # noqa: E402


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=256)
    ap.add_argument("--n-min", type=int, default=1000)
    ap.add_argument("--n-max", type=int, default=20000)
    ap.add_argument("--fa2-iters", type=int, default=1500)
    ap.add_argument("--pivots", type=int, default=50)
    ap.add_argument("--sgd-epochs", type=int, default=30)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--steps-per-graph", type=int, default=2)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=20_000)
    ap.add_argument("--out", default=W.WEIGHTS)
    args = ap.parse_args()
    log = lambda m: print(f"[train] {m}", flush=True)  # noqa: E731

    t0 = time.perf_counter()
    graphs = W.training_graphs(args.graphs, args.seed, args.n_min, args.n_max)
    sizes = [len(o) - 1 for o, _, _ in graphs]
    log(f"{len(graphs)} graphs, {sum(sizes)} nodes ({time.perf_counter() - t0:.1f}s)")
    t1 = time.perf_counter()
    fa2 = W.fa2_targets(graphs, args.fa2_iters, theta=1.0)
    log(f"size-law layouts: {args.fa2_iters} FA2 iterations on the batch ({time.perf_counter() - t1:.1f}s)")
    coef = W.fit_size_law(graphs, fa2)
    t1 = time.perf_counter()
    targets = W.stress_targets(graphs, args.pivots, args.sgd_epochs)
    log(f"shape targets: {args.sgd_epochs} SGD stress epochs per graph ({time.perf_counter() - t1:.1f}s)")
    log(f"size law log R = {coef[0]:.3f} + {coef[1]:.3f} log n + {coef[2]:.3f} log(mean degree)")
    net = jc.Mpnn(seed=args.seed)
    t2 = time.perf_counter()
    losses = W.train(net, graphs, targets, args.epochs, args.steps_per_graph, args.lr, log=log)
    log(f"trained {len(losses) * args.steps_per_graph} steps ({time.perf_counter() - t2:.1f}s), "
        f"first-epoch loss {np.mean(losses[:len(graphs)]):.4f}, last {np.mean(losses[-len(graphs):]):.4f}")
    W.save(net, coef, args.out, graphs=args.graphs, epochs=args.epochs, seed=args.seed,
           fa2_iters=args.fa2_iters, pivots=args.pivots, sgd_epochs=args.sgd_epochs,
           n_min=args.n_min, n_max=args.n_max)
    log(f"wrote {args.out}")
    print(json.dumps({"graphs": args.graphs, "nodes": int(sum(sizes)),
                      "loss_first_epoch": float(np.mean(losses[:len(graphs)])),
                      "loss_last_epoch": float(np.mean(losses[-len(graphs):])),
                      "seconds": time.perf_counter() - t0}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
