#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py

The fixtures are data (inputs' digests and expected outputs).  They pin the
oracle against drift (tests/test_golden.py re-derives every value and demands
the same bits) and give the GPU tests fixed expected outputs
(tests/test_golden_gpu.py).  Seeds and sizes follow the C1-C3 workloads at
sizes the oracle finishes in seconds.
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import mpnn_oracle as M  # noqa: E402
from oracle import oracle as O  # noqa: E402

SEED = 0xC0FFEE


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def fa2_sbm2000():
    """C1: SBM 2,000 / 4 blocks, p_in 0.02, p_out 0.001; x0 uniform in [-1,1]^3;
    theta 1.2.  Forces at x0, exact swing/traction sums, and the Morton-tree
    trajectory with exact sums after 5 and 50 iterations."""
    off, nb = O.sbm_graph(2000, 4, 0.02, 0.001, SEED)
    xyz = O.uniform_cube(2000, 1.0, SEED + 1)
    n = len(off) - 1
    deg, aoff, aadj = O.fa2_prepare(n, off, nb)
    s = O.Fa2Settings.make(theta=1.2)
    tree = O.octree_morton(xyz, deg.astype(np.float32) + 1)
    f0, stats = O.fa2_forces(xyz, None, deg, aoff, aadj, tree, s)
    x5, o5, se5, _ = O.fa2_run(off, nb, xyz, 5, s, tree_kind=1, exact=True)
    x50, o50, se50, _ = O.fa2_run(off, nb, xyz, 50, s, tree_kind=1, exact=True)
    return dict(graph_sha256=digest(off, nb), x0_sha256=digest(xyz), forces0=f0,
                sums0=O.exact_sum2(stats), pos5=x5, old5=o5, se5=se5,
                pos50_sha256=digest(x50), old50_sha256=digest(o50), se50=se50)


def octree_cube1000():
    """Octree of 1,000 bodies uniform in [-3,3]^3 with masses deg+1 of a BA(1000, 2)
    graph: sorted keys, permutation, pre-order nodes (com/mass bits, half, body,
    skip, depth); the insertion tree equals it."""
    off, nb = O.csr_from_pairs(1000, O.ba_pairs(1000, 2, 1007))
    xyz = O.uniform_cube(1000, 3.0, 1011)
    deg, _, _ = O.fa2_prepare(1000, off, nb)
    t = O.octree_morton(xyz, deg.astype(np.float32) + 1)
    N = t.N
    return dict(keys=t.keys, perm=t.perm, n_tree=np.int64(N), node_a=t.node_a[:N].copy(),
                half=t.half[:N].copy(), body=t.body[:N].copy(), skip=t.skip[:N].copy(),
                depth=t.depth[:N].copy(), x_sha256=digest(xyz), graph_sha256=digest(off, nb))


def pagerank_ba3000():
    """PageRank pull on BA(3000, 4): d 0.85, L1 tol 1e-6, at most 100 iterations."""
    off, nb = O.ba_graph(3000, 4, SEED)
    r, it, res = O.pagerank_pull(off, nb, 0.85, 100, 1e-6)
    return dict(graph_sha256=digest(off, nb), ranks=r, iters=np.int64(it), residual=np.float64(res))


def sgd_rgg2000():
    """SGD sparse stress on the largest component of a 2-D RGG (2,000 nodes, mean
    degree 8): 20 max-min pivots, u16 hops, eta from (max hop)^2 to 0.01 over 10
    sweeps, x0 uniform in [-1,1]^2 (z = 0)."""
    n0 = 2000
    xy, pairs = O.rgg_pairs(n0, math.sqrt(8.0 / (math.pi * n0)), SEED)
    off0, nb0 = O.csr_from_pairs(n0, pairs)
    _, keep = O.largest_component(n0, off0, nb0)
    sel = keep != 0xFFFFFFFF
    p = pairs.reshape(-1, 2)
    pm = keep[p]
    ok = (pm[:, 0] != 0xFFFFFFFF) & (pm[:, 1] != 0xFFFFFFFF)
    n = int(sel.sum())
    off, nb = O.csr_from_pairs(n, pm[ok].reshape(-1))
    x0 = O.uniform_cube(n, 1.0, SEED + 2).reshape(-1, 3).copy()
    x0[:, 2] = 0.0
    x0 = x0.reshape(-1)
    piv, dist = O.select_pivots(off, nb, 20, 0x5EEDC0DE12345678)
    dmax = np.float32(dist.max())
    eta_max = float(dmax * dmax)
    pos = x0.copy()
    for t in range(10):
        pos = O.sgd_sweep(piv, dist, pos, O.anneal_eta(t, eta_max, 0.01, 10))
    return dict(offsets=off, neighbors=nb, x0=x0, pivots=piv, dist_sha256=digest(dist),
                pos10=pos)


MPNN_SEED = 5
GRAD_STRIDE = 97


def mpnn_inputs():
    """BA(40, 2) with parameters, features and target drawn from PCG64(5)."""
    off, nb = O.ba_graph(40, 2, 3)
    rng = np.random.default_rng(MPNN_SEED)
    P = sum((k + 1) * M.HID for k in M.K) + (M.HID + 1) * 3
    params = rng.standard_normal(P) * 0.1
    x = rng.standard_normal((40, 16)) * 0.5
    tgt = rng.standard_normal(3 * 40)
    return off, nb, params, x, tgt


def mpnn_tiny():
    """The warm-start network in float64 (no bf16 rounding): forward output, and
    loss + every GRAD_STRIDE-th gradient entry against the seeded target with
    anchors {0, 7}.  The parameters are regenerated from the seed, not stored."""
    off, nb, params, x, tgt = mpnn_inputs()
    y, _ = M.forward(params, off, nb, x, emulate_bf16=False)
    L, g = M.loss_grad(params, off, nb, x, tgt, [0, 7], emulate_bf16=False)
    return dict(y=y, loss=np.float64(L), grad_sample=g[::GRAD_STRIDE], n_params=np.int64(len(params)))


def generators():
    """Digests of the seeded workload graphs (C1-C3) and SplitMix64 outputs."""
    out = {}
    off, nb = O.ba_graph(1_000_000, 4, SEED)
    out["ba_1000000_4"] = {"n": 1_000_000, "nnz": int(len(nb)), "sha256": digest(off, nb)}
    off, nb = O.sbm_graph(2000, 4, 0.02, 0.001, SEED)
    out["sbm_2000_4"] = {"n": 2000, "nnz": int(len(nb)), "sha256": digest(off, nb)}
    out["cube_1000000_10_seed+1"] = digest(O.uniform_cube(1_000_000, 10.0, SEED + 1))
    out["splitmix_0_first4"] = [int(v) for v in O.splitmix_u64(0, 4)]
    return out


FIXTURES = {
    "fa2_sbm2000.npz": fa2_sbm2000,
    "octree_cube1000.npz": octree_cube1000,
    "pagerank_ba3000.npz": pagerank_ba3000,
    "sgd_rgg2000.npz": sgd_rgg2000,
    "mpnn_tiny.npz": mpnn_tiny,
}


def main() -> int:
    for name, fn in FIXTURES.items():
        data = fn()
        np.savez_compressed(os.path.join(HERE, name), **{k: np.asarray(v) for k, v in data.items()})
        print(f"wrote {name}")
    with open(os.path.join(HERE, "generators.json"), "w") as fh:
        json.dump(generators(), fh, indent=1)
    print("wrote generators.json")
    return 0


if __name__ == "__main__":
    sys.exit(main())
