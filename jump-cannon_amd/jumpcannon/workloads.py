"""Seeded synthetic workloads of the benchmark configurations.

All inputs come from SplitMix64 (crates/graph-compute/src/stability.rs:11-26)
through the library's own generators; seeds are fixed here so tests, the
benchmark and the fixtures agree.
"""
from __future__ import annotations

import math

import numpy as np

from . import ba_graph, rgg_pairs, csr_from_pairs, sbm_graph, uniform_cube

SEED = 0xC0FFEE


def sbm_2000():
    """SBM, 2,000 nodes in 4 equal blocks, p_in 0.02, p_out 0.001; positions
    uniform in [-1,1]^3.  FA2+BH theta 1.2, scaling 2, gravity 1, 500 iterations."""
    off, nb = sbm_graph(2000, 4, 0.02, 0.001, SEED)
    xyz = uniform_cube(2000, 1.0, SEED + 1)
    return off, nb, xyz


def ba_1m(n: int = 1_000_000, m: int = 4):
    """Barabasi-Albert (synthetic_scale_free) n=1,000,000, m=4; positions uniform
    in [-10,10]^3.  FA2+BH theta 0.8, 200 iterations; PageRank on the same CSR."""
    off, nb = ba_graph(n, m, SEED)
    xyz = uniform_cube(n, 10.0, SEED + 1)
    return off, nb, xyz


def rgg_200k(n: int = 200_000, mean_degree: float = 8.0):
    """2-D random geometric graph in the unit square, radius for the given mean
    degree, largest component only (ids relabelled in increasing order)."""
    r = math.sqrt(mean_degree / (math.pi * n))
    xy, pairs = rgg_pairs(n, r, SEED)
    off, nb = csr_from_pairs(n, pairs)
    keep = largest_component_map(off, nb)
    sel = keep != 0xFFFFFFFF
    p = pairs.reshape(-1, 2)
    pm = keep[p]
    ok = (pm[:, 0] != 0xFFFFFFFF) & (pm[:, 1] != 0xFFFFFFFF)
    n2 = int(sel.sum())
    off2, nb2 = csr_from_pairs(n2, pm[ok].reshape(-1))
    return off2, nb2, xy.reshape(-1, 2)[sel]


def batch_4096(n_graphs: int = 4096, seed: int = SEED + 100):
    """Graphs of size n = floor(50 * 400^u), u uniform in [0,1) (log-uniform over
    50..20,000).  Even indices: SBM with B = 2 + (h mod 7) equal blocks,
    p_in = 8/n_block, p_out = 0.5/n.  Odd indices: BA with m = 2 + (h mod 4).
    u and h come from SplitMix64(seed + g); positions uniform in [-1,1]^3.
    FA2+BH theta 1.0, 500 iterations."""
    from . import _splitmix_pair

    graphs, pos = [], []
    for g in range(n_graphs):
        u, h = _splitmix_pair(seed + g)
        n = int(50.0 * 400.0 ** u)
        if g % 2 == 0:
            B = 2 + (h % 7)
            nblock = n / B
            off, nb = sbm_graph(n, B, min(1.0, 8.0 / nblock), 0.5 / n, seed + g)
        else:
            off, nb = ba_graph(n, 2 + (h % 4), seed + g)
        graphs.append((off, nb))
        pos.append(uniform_cube(len(off) - 1, 1.0, seed + 50_000 + g))
    return graphs, pos


def plane_init(n: int, seed: int = SEED + 2):
    """2-D layout seed for the stress workloads: x, y = 2u - 1 in [-1,1], z = 0."""
    xyz = uniform_cube(n, 1.0, seed).reshape(-1, 3).copy()
    xyz[:, 2] = 0.0
    return xyz.reshape(-1)


def largest_component_map(off, nb):
    """keep[v] = new id in the largest component (ties: the component holding
    the smallest node id), else 0xFFFFFFFF."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components

    n = len(off) - 1
    if n == 0:
        return np.zeros(0, np.uint32)
    a = csr_matrix((np.ones(len(nb), np.int8), np.asarray(nb, np.int64), np.asarray(off, np.int64)),
                   shape=(n, n))
    _, lab = connected_components(a, directed=False)
    sizes = np.bincount(lab)
    # components are labelled in order of their smallest node, so argmax breaks
    # ties towards the smallest node id
    best = int(np.argmax(sizes))
    keep = np.full(n, 0xFFFFFFFF, np.uint32)
    idx = np.nonzero(lab == best)[0]
    keep[idx] = np.arange(len(idx), dtype=np.uint32)
    return keep
