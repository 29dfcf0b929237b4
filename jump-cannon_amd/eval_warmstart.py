#!/usr/bin/env python3
"""Evaluates the committed warm-start network on held-out graphs (seeds and
sizes not in the training set) and prints one JSON line per graph.

    python jump-cannon_amd/eval_warmstart.py [--sizes 20000 50000 100000] [--horizon 2000]

For every size, one Barabasi-Albert graph (m=4) and one stochastic block model
(4 blocks, p_in = 8/block size, p_out = 0.5/n), positions uniform in
[-10,10]^3, FA2+BH theta 0.8, judged as warmstart.evaluate does: time to stress
<= 1.01 x the plain run's converged stress S*, and time to both stress and
edge-length CV within 1 % of the plain run's converged values."""
from __future__ import annotations

import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import jumpcannon as jc  # noqa: E402
from jumpcannon import warmstart as W  # noqa: E402

KEEP = ("plain_init_stress", "plain_stress_final", "learned_init_stress", "learned_stress_final",
        "plain_init_cv", "plain_cv_final", "learned_init_cv", "learned_cv_final", "plain_iters_to_target",
        "learned_iters_to_target", "plain_ms_to_target", "learned_ms_to_target", "speedup",
        "learned_max_stress_after_target", "plain_iters_to_both", "learned_iters_to_both", "plain_ms_to_both",
        "learned_ms_to_both", "speedup_both", "learned_stays_within_both", "fallback_to_plain", "inference_ms")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", type=int, nargs="+", default=[20_000, 50_000, 100_000])
    ap.add_argument("--horizon", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=0x4E4F54)  # not a training seed (those are 20,000 + g)
    args = ap.parse_args()
    net, coef = W.load()
    for n in args.sizes:
        for kind in ("ba", "sbm"):
            s = args.seed + n + (0 if kind == "ba" else 7)
            off, nb = jc.ba_graph(n, 4, s) if kind == "ba" else jc.sbm_graph(n, 4, min(1.0, 8.0 / (n / 4)), 0.5 / n, s)
            x0 = jc.uniform_cube(len(off) - 1, W.INIT_HALF, s + 1)
            r = W.evaluate(net, coef, off, nb, x0, horizon=args.horizon, every=50, theta=0.8, ablation=False,
                           multilevel=False)
            out = {"graph": kind, "n": len(off) - 1, "nnz": len(nb)}
            out.update({k: (round(r[k], 5) if isinstance(r[k], float) else r[k]) for k in KEEP})
            print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
