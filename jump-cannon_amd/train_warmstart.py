#!/usr/bin/env python3
"""Trains the FA2 warm-start network on the GPU and writes its weights.

    python jump-cannon_amd/train_warmstart.py [--graphs 20000] [--n-min 1000] [--n-max 100000]
        [--stop G --checkpoint PATH] [--resume PATH] [--out PATH]

Training set: a seeded stream of SBM/BA graphs of log-uniform size in
[n_min, n_max], each with a random initial layout x0 in [-10,10]^3.  Every
graph is generated, given its target, trained on and dropped, so the stream's
length does not bound memory.  Shape targets are the layouts the SGD
sparse-stress engine reaches from x0; the size law comes from the plain FA2+BH
layouts of the first --law-graphs graphs after --fa2-iters iterations (all in
one batched run).  Every forward, backward and Adam step runs in the library's
own HIP kernels; the learning rate follows a cosine from --lr to --lr-min over
the whole stream.

A run may stop after graph --stop and write a checkpoint (parameters, Adam
moments and step, size law, losses); --resume continues from it exactly.  The
default output is the file the warm start loads
(jumpcannon/warmstart_weights.npz), written once the stream is finished.
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
    ap.add_argument("--graphs", type=int, default=20_000)
    ap.add_argument("--n-min", type=int, default=1000)
    ap.add_argument("--n-max", type=int, default=100_000)
    ap.add_argument("--law-graphs", type=int, default=256)
    ap.add_argument("--fa2-iters", type=int, default=1500)
    ap.add_argument("--pivots", type=int, default=50)
    ap.add_argument("--sgd-epochs", type=int, default=30)
    ap.add_argument("--steps-per-graph", type=int, default=2)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-min", type=float, default=2e-4)
    ap.add_argument("--seed", type=int, default=20_000)
    ap.add_argument("--stop", type=int, default=None, help="stop after this many graphs and checkpoint")
    ap.add_argument("--checkpoint", default=None)
    ap.add_argument("--resume", default=None)
    ap.add_argument("--out", default=W.WEIGHTS)
    args = ap.parse_args()
    log = lambda m: print(f"[train] {m}", flush=True)  # noqa: E731
    t0 = time.perf_counter()
    net = jc.Mpnn(seed=args.seed)
    if args.resume:
        with np.load(args.resume, allow_pickle=False) as z:
            net.set_params(z["params"])
            net.set_adam_state(z["adam_m"], z["adam_v"], int(z["adam_step"]))
            coef = z["size_law"].copy()
            losses = list(z["losses"])
            start = int(z["next_graph"])
            step = int(z["adam_step"])
            if int(z["graphs"]) != args.graphs or int(z["seed"]) != args.seed:
                raise SystemExit("checkpoint is from a different stream (--graphs/--seed)")
        log(f"resumed at graph {start} (Adam step {step})")
    else:
        laws = W.training_graphs(args.law_graphs, args.seed, args.n_min, args.n_max)
        log(f"size law: {len(laws)} graphs, {sum(len(o) - 1 for o, _, _ in laws)} nodes")
        fa2 = W.fa2_targets(laws, args.fa2_iters, theta=1.0)
        coef = W.fit_size_law(laws, fa2)
        del laws, fa2
        log(f"size law log R = {coef[0]:.3f} + {coef[1]:.3f} log n + {coef[2]:.3f} log(mean degree) "
            f"({time.perf_counter() - t0:.1f}s)")
        losses, start = [], 0
    stop = min(args.graphs, args.stop if args.stop is not None else args.graphs)
    t1 = time.perf_counter()
    losses += W.train_stream(net, start, stop, args.graphs, args.seed, args.n_min, args.n_max,
                             args.steps_per_graph, args.lr, args.lr_min, pivots=args.pivots,
                             sgd_epochs=args.sgd_epochs, log=log)
    log(f"graphs {start}..{stop} in {time.perf_counter() - t1:.1f}s")
    if stop < args.graphs:
        if not args.checkpoint:
            raise SystemExit("--stop before the end of the stream needs --checkpoint")
        m1, m2, step = net.adam_state()
        np.savez(args.checkpoint, params=net.params(), adam_m=m1, adam_v=m2, adam_step=np.uint64(step),
                 size_law=np.asarray(coef, np.float64), losses=np.asarray(losses, np.float64),
                 next_graph=np.int64(stop), graphs=np.int64(args.graphs), seed=np.int64(args.seed))
        log(f"checkpoint at graph {stop}: {args.checkpoint}")
        return 0
    W.save(net, coef, args.out, graphs=args.graphs, passes=1, seed=args.seed, fa2_iters=args.fa2_iters,
           law_graphs=args.law_graphs, steps_per_graph=args.steps_per_graph, pivots=args.pivots,
           sgd_epochs=args.sgd_epochs, n_min=args.n_min, n_max=args.n_max)
    log(f"wrote {args.out}")
    k = min(500, len(losses))
    print(json.dumps({"graphs": args.graphs, "loss_first_500": float(np.mean(losses[:k])),
                      "loss_last_500": float(np.mean(losses[-k:])), "seconds_this_run": time.perf_counter() - t0}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
