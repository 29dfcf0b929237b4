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


def largest_component_map(off, nb):
    """keep[v] = new id in the largest component (ties: smallest node id), else ~0."""
    n = len(off) - 1
    comp = np.full(n, -1, np.int64)
    best, best_size, c = -1, 0, 0
    off = np.asarray(off, np.int64)
    for s in range(n):
        if comp[s] >= 0:
            continue
        stack = [s]
        comp[s] = c
        size = 0
        while stack:
            v = stack.pop()
            size += 1
            for u in nb[off[v]:off[v + 1]]:
                if comp[u] < 0:
                    comp[u] = c
                    stack.append(int(u))
        if size > best_size:
            best, best_size = c, size
        c += 1
    keep = np.full(n, 0xFFFFFFFF, np.uint32)
    idx = np.nonzero(comp == best)[0]
    keep[idx] = np.arange(len(idx), dtype=np.uint32)
    return keep
