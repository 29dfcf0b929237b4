"""ctypes front-end of the CPU oracle (libjc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``, always as the checker or the timed
CPU baseline, never as the product path.  The C sources cite the reference
(ocasazza/jump-cannon) file:line each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjc_oracle.so")

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

INTERNAL = 0xFFFFFFFF
END = 0xFFFFFFFF
LEVELS = 21


class Fa2Settings(C.Structure):
    """Field-for-field Fa2BhSettings (crates/graph-compute/src/engines/fa2_bh.rs:85-125)."""

    _fields_ = [
        ("gravity", C.c_float),
        ("strong_gravity", C.c_uint32),
        ("scaling_ratio", C.c_float),
        ("edge_weight_influence", C.c_float),
        ("jitter_tolerance", C.c_float),
        ("lin_log_mode", C.c_uint32),
        ("prevent_overlap", C.c_uint32),
        ("time_step", C.c_float),
        ("theta", C.c_float),
        ("max_displacement", C.c_float),
    ]

    @classmethod
    def make(cls, **kw):
        d = dict(gravity=1.0, strong_gravity=0, scaling_ratio=2.0, edge_weight_influence=1.0,
                 jitter_tolerance=1.0, lin_log_mode=0, prevent_overlap=0, time_step=1.0,
                 theta=0.7, max_displacement=10.0)
        d.update(kw)
        return cls(**d)


def build() -> str:
    subprocess.check_call(["make", "-s", "-C", HERE])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    S = C.POINTER(Fa2Settings)
    sig = {
        "jco_next_u64": (C.c_uint64, [C.POINTER(C.c_uint64)]),
        "jco_next_f32": (C.c_float, [C.POINTER(C.c_uint64)]),
        "jco_uniform_cube": (None, [C.c_uint32, C.c_float, C.c_uint64, f32p]),
        "jco_ball_seed": (None, [C.c_uint32, C.c_uint64, f32p]),
        "jco_ba_pairs": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint64, u32p, C.c_uint64]),
        "jco_sbm_pairs": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_uint64,
                                       u32p, C.c_uint64]),
        "jco_rgg_pairs": (C.c_uint64, [C.c_uint32, C.c_double, C.c_uint64, f32p, u32p, C.c_uint64]),
        "jco_csr_from_pairs": (C.c_uint64, [C.c_uint32, u32p, C.c_uint64, u32p, u32p]),
        "jco_largest_component": (C.c_uint32, [C.c_uint32, u32p, u32p, u32p]),
        "jco_fa2_prepare": (C.c_uint64, [C.c_uint32, u32p, u32p, u32p, u32p, u32p]),
        "jco_octree_ref": (C.c_uint32, [C.c_uint32, f32p, f32p, C.c_uint32, f32p, f32p, u32p, u32p,
                                        u32p, f32p, f32p]),
        "jco_octree_morton": (C.c_uint32, [C.c_uint32, f32p, f32p, f32p, f32p, u32p, u32p, u32p,
                                           u64p, u32p, u32p, f32p]),
        "jco_morton_keys": (None, [C.c_uint32, f32p, u64p, f32p]),
        "jco_fa2_forces": (None, [C.c_uint32, f32p, f32p, u32p, u32p, u32p, C.c_uint32, f32p, f32p,
                                  u32p, u32p, S, f32p, f32p, C.c_int]),
        "jco_fa2_forces_f64": (None, [C.c_uint32, f32p, u32p, u32p, u32p, C.c_uint32, f32p, f32p,
                                      u32p, u32p, S, f64p, u64p, C.c_int]),
        "jco_fa2_brute_forces": (None, [C.c_uint32, f32p, f32p, u32p, u32p, u32p, S, f32p, f32p,
                                        C.c_int]),
        "jco_fa2_brute_f64": (None, [C.c_uint32, f32p, u32p, u32p, u32p, S, f64p]),
        "jco_exact_sum2": (None, [C.c_uint32, f32p, f64p]),
        "jco_speed_update": (C.c_float, [f32p, C.c_uint32, C.c_float, f32p, C.c_int, f64p]),
        "jco_fa2_apply": (None, [C.c_uint32, f32p, f32p, f32p, f32p, u32p, S, C.c_float, C.c_int]),
        "jco_fa2_run": (C.c_int, [C.c_uint32, u32p, u32p, f32p, f32p, f32p, C.c_uint32, S, C.c_int,
                                  C.c_int, C.c_int, f64p]),
        "jco_pagerank_pull": (C.c_int, [C.c_uint32, u32p, u32p, C.c_float, C.c_uint32, C.c_double,
                                        f32p, C.POINTER(C.c_uint32), C.POINTER(C.c_double), C.c_int]),
        "jco_pagerank_push": (None, [C.c_uint32, u32p, u32p, C.c_float, C.c_uint32, f32p]),
        "jco_pagerank_f64": (None, [C.c_uint32, u32p, u32p, C.c_double, C.c_uint32, f64p]),
        "jco_spmv": (None, [C.c_uint32, u32p, u32p, f32p, f32p, f32p]),
        "jco_spmv_hybrid": (None, [C.c_uint32, u32p, u32p, f32p, f32p, f32p]),
        "jco_spmv_f16": (None, [C.c_uint32, u32p, u32p, f32p, f32p, f32p]),
        "jco_f32_to_f16": (C.c_uint16, [C.c_float]),
        "jco_f16_to_f32": (C.c_float, [C.c_uint16]),
        "jco_bfs": (None, [C.c_uint32, u32p, u32p, C.c_uint32, u32p]),
        "jco_connected_components": (None, [C.c_uint32, u32p, u32p, u32p]),
        "jco_contract_level": (C.c_uint32, [C.c_uint32, u32p, C.c_uint64, u32p, u32p, u64p]),
        "jco_fold_up_level": (None, [C.c_uint32, f32p, u32p, C.c_uint32, f32p]),
        "jco_prolong": (None, [f32p, u32p, C.c_uint32, C.c_float, C.c_uint32, f32p]),
        "jco_level_csr": (None, [C.c_uint32, u32p, C.c_uint64, u32p, u32p]),
        "jco_select_pivots": (C.c_uint32, [C.c_uint32, u32p, u32p, C.c_uint32, C.c_uint64, u32p,
                                           u16p]),
        "jco_anneal_eta": (C.c_float, [C.c_uint64, C.c_float, C.c_float, C.c_uint32]),
        "jco_sgd_sweep": (None, [C.c_uint32, C.c_uint32, u32p, u16p, f32p, f32p, C.c_float, C.c_int]),
        "jco_sgd_cpu_run": (C.c_uint32, [C.c_uint32, u32p, u32p, f32p, C.c_uint32, C.c_uint32, C.c_float,
                                         C.c_float, C.c_uint32, C.c_uint64, C.c_uint32]),
        "jco_stress_normalized": (C.c_double, [f32p, u32p, u32p, f32p, C.c_uint64]),
        "jco_edge_length_cv": (C.c_double, [f32p, u32p, u32p, C.c_uint64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ---------------------------------------------------------------- fixtures
def splitmix_u64(seed: int, count: int) -> np.ndarray:
    st = C.c_uint64(seed)
    return np.array([lib().jco_next_u64(C.byref(st)) for _ in range(count)], dtype=np.uint64)


def uniform_cube(n: int, half: float, seed: int) -> np.ndarray:
    out = np.empty(3 * n, dtype=np.float32)
    lib().jco_uniform_cube(n, half, seed, out)
    return out


def ball_seed(n: int, seed: int) -> np.ndarray:
    out = np.empty(3 * n, dtype=np.float32)
    lib().jco_ball_seed(n, seed, out)
    return out


def csr_from_pairs(n: int, pairs: np.ndarray):
    pairs = _u32(pairs).reshape(-1)
    offsets = np.empty(n + 1, dtype=np.uint32)
    nbrs = np.empty(max(len(pairs), 1), dtype=np.uint32)
    nnz = lib().jco_csr_from_pairs(n, pairs, len(pairs) // 2, offsets, nbrs)
    return offsets, nbrs[:nnz].copy()


def ba_pairs(n: int, m: int, seed: int) -> np.ndarray:
    n2 = max(n, 2)
    cap = 1 + (n2 - 2) * max(m, 1)
    pairs = np.empty(2 * cap, dtype=np.uint32)
    np_ = lib().jco_ba_pairs(n, m, seed, pairs, cap)
    return pairs[: 2 * np_]


def ba_graph(n: int, m: int, seed: int):
    """synthetic_scale_free (crates/graph-compute/src/stability.rs:57-82)."""
    return csr_from_pairs(max(n, 2), ba_pairs(n, m, seed))


def sbm_pairs(n: int, blocks: int, p_in: float, p_out: float, seed: int) -> np.ndarray:
    cap = 1 << 16
    while True:
        pairs = np.empty(2 * cap, dtype=np.uint32)
        cnt = lib().jco_sbm_pairs(n, blocks, p_in, p_out, seed, pairs, cap)
        if cnt <= cap:
            return pairs[: 2 * cnt]
        cap = int(cnt)


def sbm_graph(n, blocks, p_in, p_out, seed):
    return csr_from_pairs(n, sbm_pairs(n, blocks, p_in, p_out, seed))


def rgg_pairs(n: int, radius: float, seed: int):
    xy = np.empty(2 * n, dtype=np.float32)
    cap = max(16, int(n * 6))
    while True:
        pairs = np.empty(2 * cap, dtype=np.uint32)
        cnt = lib().jco_rgg_pairs(n, radius, seed, xy, pairs, cap)
        if cnt <= cap:
            return xy, pairs[: 2 * cnt]
        cap = int(cnt)


def largest_component(n, offsets, nbrs):
    keep = np.empty(n, dtype=np.uint32)
    size = lib().jco_largest_component(n, _u32(offsets), _u32(nbrs), keep)
    return size, keep


# ---------------------------------------------------------------- FA2
def fa2_prepare(n, offsets, nbrs):
    deg = np.empty(max(n, 1), dtype=np.uint32)
    att_off = np.empty(n + 1, dtype=np.uint32)
    att_adj = np.empty(max(len(nbrs), 1), dtype=np.uint32)
    m = lib().jco_fa2_prepare(n, _u32(offsets), _u32(nbrs), deg, att_off, att_adj)
    return deg[:n], att_off, att_adj[:m] if m else np.zeros(1, np.uint32)


class Tree:
    def __init__(self, N, node_a, half, body, skip, depth, root4, keys=None, perm=None, n_valid=None):
        self.N = N
        self.node_a = node_a[: 4 * N].reshape(-1, 4) if N else np.zeros((0, 4), np.float32)
        self.node_a_flat = node_a
        self.half = half
        self.body = body
        self.skip = skip
        self.depth = depth
        self.root4 = root4
        self.keys = keys
        self.perm = perm
        self.n_valid = n_valid


def octree_ref(xyz, mass, max_nodes=None):
    xyz = _f32(xyz)
    n = len(xyz) // 3
    cap = max_nodes if max_nodes is not None else 2 * n + 8
    na = np.zeros(4 * max(cap, 1), np.float32)
    half = np.zeros(max(cap, 1), np.float32)
    body = np.zeros(max(cap, 1), np.uint32)
    skip = np.zeros(max(cap, 1), np.uint32)
    depth = np.zeros(max(cap, 1), np.uint32)
    root4 = np.zeros(4, np.float32)
    centre = np.zeros(3 * max(cap, 1), np.float32)
    N = lib().jco_octree_ref(n, xyz, _f32(mass), cap, na, half, body, skip, depth, root4, centre)
    t = Tree(N, na, half, body, skip, depth, root4)
    t.centre = centre[: 3 * N].reshape(-1, 3)
    return t


def octree_morton(xyz, mass):
    xyz = _f32(xyz)
    n = len(xyz) // 3
    cap = (LEVELS + 2) * n + 8
    na = np.zeros(4 * cap, np.float32)
    half = np.zeros(cap, np.float32)
    body = np.zeros(cap, np.uint32)
    skip = np.zeros(cap, np.uint32)
    depth = np.zeros(cap, np.uint32)
    keys = np.zeros(max(n, 1), np.uint64)
    perm = np.zeros(max(n, 1), np.uint32)
    nv = C.c_uint32(0)
    root4 = np.zeros(4, np.float32)
    nvarr = np.zeros(1, np.uint32)
    N = lib().jco_octree_morton(n, xyz, _f32(mass), na, half, body, skip, depth, keys, perm, nvarr,
                                root4)
    return Tree(N, na, half, body, skip, depth, root4, keys[:n], perm[:n], int(nvarr[0]))


def morton_keys(xyz):
    xyz = _f32(xyz)
    n = len(xyz) // 3
    keys = np.zeros(max(n, 1), np.uint64)
    root4 = np.zeros(4, np.float32)
    lib().jco_morton_keys(n, xyz, keys, root4)
    return keys[:n], root4


def fa2_forces(xyz, old_force, deg, att_off, att_adj, tree: Tree, settings: Fa2Settings, threads=8):
    xyz = _f32(xyz)
    n = len(xyz) // 3
    force = np.zeros(3 * n, np.float32)
    stats = np.zeros(2 * max(n, 1), np.float32)
    old = _f32(old_force) if old_force is not None else np.zeros(3 * n, np.float32)
    lib().jco_fa2_forces(n, xyz, old, _u32(deg), _u32(att_off), _u32(att_adj), tree.N,
                         tree.node_a_flat, tree.half, tree.body, tree.skip, C.byref(settings),
                         force, stats, threads)
    return force, stats[: 2 * n]


def fa2_forces_f64(xyz, deg, att_off, att_adj, tree: Tree, settings: Fa2Settings, threads=8):
    xyz = _f32(xyz)
    n = len(xyz) // 3
    force = np.zeros(3 * n, np.float64)
    flips = np.zeros(1, np.uint64)
    lib().jco_fa2_forces_f64(n, xyz, _u32(deg), _u32(att_off), _u32(att_adj), tree.N,
                             tree.node_a_flat, tree.half, tree.body, tree.skip, C.byref(settings),
                             force, flips, threads)
    return force, int(flips[0])


def fa2_brute_forces(xyz, old_force, deg, att_off, att_adj, settings: Fa2Settings, threads=8):
    """The brute-force engine's f32 force pass (force_atlas2.wgsl:69-135):
    returns (force 3n, stats 2n)."""
    xyz = _f32(xyz)
    n = len(xyz) // 3
    force = np.zeros(3 * n, np.float32)
    stats = np.zeros(2 * max(n, 1), np.float32)
    old = _f32(old_force) if old_force is not None else np.zeros(3 * n, np.float32)
    lib().jco_fa2_brute_forces(n, xyz, old, _u32(deg), _u32(att_off), _u32(att_adj),
                               C.byref(settings), force, stats, threads)
    return force, stats[: 2 * n]


def fa2_brute_f64(xyz, deg, att_off, att_adj, settings: Fa2Settings):
    xyz = _f32(xyz)
    n = len(xyz) // 3
    force = np.zeros(3 * n, np.float64)
    lib().jco_fa2_brute_f64(n, xyz, _u32(deg), _u32(att_off), _u32(att_adj), C.byref(settings), force)
    return force


def exact_sum2(stats):
    stats = _f32(stats)
    out = np.zeros(2, np.float64)
    lib().jco_exact_sum2(len(stats) // 2, stats, out)
    return out


def speed_update(se, n, jt, stats, exact=False):
    se = np.ascontiguousarray(se, dtype=np.float32)
    sums = np.zeros(2, np.float64)
    lib().jco_speed_update(se, n, jt, _f32(stats), 1 if exact else 0, sums)
    return se, sums


def fa2_apply(xyz, force, old_force, deg, settings: Fa2Settings, speed: float):
    """One fa2_apply pass (fa2_barnes_hut.wgsl:203-226).  Returns (new xyz,
    new F_{t-1} = force)."""
    xyz = _f32(xyz)
    n = len(xyz) // 3
    out = np.empty_like(xyz)
    old = _f32(old_force).copy()
    lib().jco_fa2_apply(n, xyz, out, _f32(force), old, _u32(deg), C.byref(settings), speed, 1)
    return out, old


def fa2_run(offsets, nbrs, xyz, iters, settings: Fa2Settings, tree_kind=1, exact=True, threads=8,
            old_force=None, se=None):
    """Whole FA2 iterations (fa2_bh.rs:546-692; tree_kind 2: the brute-force
    engine's, fa2_brute.rs:454-565).  Returns (xyz, old_force, se, timings)."""
    xyz = _f32(xyz).copy()
    n = len(xyz) // 3
    old = np.zeros(3 * n, np.float32) if old_force is None else _f32(old_force).copy()
    se = np.array([1.0, 1.0], np.float32) if se is None else np.array(se, np.float32)
    t = np.zeros(3, np.float64)
    lib().jco_fa2_run(n, _u32(offsets), _u32(nbrs), xyz, old, se, iters, C.byref(settings),
                      tree_kind, 1 if exact else 0, threads, t)
    return xyz, old, se, t


# ---------------------------------------------------------------- PageRank
def pagerank_pull(offsets, nbrs, damping=0.85, max_iters=100, tol=1e-6, threads=8):
    n = len(offsets) - 1
    ranks = np.zeros(max(n, 1), np.float32)
    it = C.c_uint32(0)
    res = C.c_double(0)
    rc = lib().jco_pagerank_pull(n, _u32(offsets), _u32(nbrs) if len(nbrs) else np.zeros(1, np.uint32),
                                 damping, max_iters, tol, ranks, C.byref(it), C.byref(res), threads)
    if rc != 0:
        raise ValueError("dangling node")
    return ranks[:n], it.value, res.value


def _spmv(fn, offsets, nbrs, weights, x):
    n = len(offsets) - 1
    y = np.zeros(max(n, 1), np.float32)
    pad_u = np.zeros(1, np.uint32)
    pad_f = np.zeros(1, np.float32)
    fn(n, _u32(offsets), _u32(nbrs) if len(nbrs) else pad_u,
       _f32(weights) if len(weights) else pad_f, _f32(x) if len(x) else pad_f, y)
    return y[:n]


def spmv(offsets, nbrs, weights, x):
    """cpu_spmv (spmv.rs:88-99): each row summed in CSR order in f32."""
    return _spmv(lib().jco_spmv, offsets, nbrs, weights, x)


def spmv_hybrid(offsets, nbrs, weights, x):
    """gpu_spmv_hybrid's order (spmv_hybrid.wgsl:44-84): 64 lane-strided partials
    and a stride-halving tree."""
    return _spmv(lib().jco_spmv_hybrid, offsets, nbrs, weights, x)


def spmv_f16(offsets, nbrs, weights, x):
    """gpu_spmv_f16 (spmv_f16.wgsl:27-45): f16-rounded weights and x, f32 sums."""
    return _spmv(lib().jco_spmv_f16, offsets, nbrs, weights, x)


def f32_to_f16_bits(values):
    """Nearest-even binary16 bits of each f32 (pack_f16_pairs, spmv.rs:68-86)."""
    f = lib().jco_f32_to_f16
    return np.array([f(float(v)) for v in np.asarray(values, np.float32)], np.uint16)


def f16_bits_to_f32(bits):
    f = lib().jco_f16_to_f32
    return np.array([f(int(b)) for b in np.asarray(bits, np.uint16)], np.float32)


def pagerank_push(offsets, nbrs, damping=0.85, iters=50):
    n = len(offsets) - 1
    ranks = np.zeros(max(n, 1), np.float32)
    lib().jco_pagerank_push(n, _u32(offsets), _u32(nbrs) if len(nbrs) else np.zeros(1, np.uint32),
                            damping, iters, ranks)
    return ranks[:n]


def pagerank_f64(offsets, nbrs, damping=0.85, iters=50):
    n = len(offsets) - 1
    ranks = np.zeros(max(n, 1), np.float64)
    lib().jco_pagerank_f64(n, _u32(offsets), _u32(nbrs) if len(nbrs) else np.zeros(1, np.uint32),
                           damping, iters, ranks)
    return ranks[:n]


# ---------------------------------------------------------------- SGD
def bfs(offsets, nbrs, src):
    n = len(offsets) - 1
    d = np.zeros(max(n, 1), np.uint32)
    lib().jco_bfs(n, _u32(offsets), _u32(nbrs) if len(nbrs) else np.zeros(1, np.uint32), src, d)
    return d[:n]


def connected_components(offsets, nbrs):
    n = len(offsets) - 1
    lab = np.zeros(max(n, 1), np.uint32)
    lib().jco_connected_components(n, _u32(offsets), _u32(nbrs) if len(nbrs) else np.zeros(1, np.uint32),
                                   lab)
    return lab[:n]


def select_pivots(offsets, nbrs, k, seed=0x5EEDC0DE12345678):
    n = len(offsets) - 1
    kk = max(1, min(k, n))
    piv = np.zeros(kk, np.uint32)
    dist = np.zeros(max(n * kk, 1), np.uint16)
    got = lib().jco_select_pivots(n, _u32(offsets), _u32(nbrs) if len(nbrs) else np.zeros(1, np.uint32),
                                  k, seed, piv, dist)
    if got == 0xFFFFFFFF:
        raise ValueError("hop distance exceeds u16")
    return piv, dist[: n * kk]


def anneal_eta(sweep, eta_max, eta_min, n_anneal):
    return lib().jco_anneal_eta(sweep, eta_max, eta_min, n_anneal)


def sgd_sweep(pivots, dist, pos, eta, threads=8):
    pos = _f32(pos)
    n = len(pos) // 3
    out = np.empty_like(pos)
    lib().jco_sgd_sweep(n, len(pivots), _u32(pivots), np.ascontiguousarray(dist, np.uint16), pos, out,
                        eta, threads)
    return out


def sgd_cpu_run(offsets, nbrs, xyz, steps, n_pivots=50, sweeps_per_step=1, eta_max=1.0, eta_min=0.01,
                n_anneal=30, seed=0x5EEDC0DE12345678):
    """The reference CPU s_gd2 engine (SgdStressEngine, sgd_stress.rs:251-316):
    `steps` steps from xyz.  Returns the positions (3n)."""
    pos = _f32(xyz).copy()
    n = len(offsets) - 1
    k = lib().jco_sgd_cpu_run(n, _u32(offsets), _u32(nbrs) if len(nbrs) else np.zeros(1, np.uint32), pos,
                              n_pivots, sweeps_per_step, eta_max, eta_min, n_anneal, seed, steps)
    if k == 0xFFFFFFFF:
        raise ValueError("hop distance exceeds u16")
    return pos


def all_pairs_terms(offsets, nbrs):
    """(i, j, d) for every reachable pair i < j with hop distance d > 0: the
    term set of solved_layouts.rs scale_normalized_stress (all_pairs BFS)."""
    n = len(offsets) - 1
    ti, tj, td = [], [], []
    for s in range(n):
        d = bfs(offsets, nbrs, s)
        j = np.arange(s + 1, n)
        dj = d[s + 1:]
        ok = (dj != 0xFFFFFFFF) & (dj != 0)
        ti.append(np.full(ok.sum(), s, np.uint32))
        tj.append(j[ok].astype(np.uint32))
        td.append(dj[ok].astype(np.float32))
    cat = lambda a, t: np.concatenate(a).astype(t) if a else np.zeros(0, t)
    return cat(ti, np.uint32), cat(tj, np.uint32), cat(td, np.float32)


# ---------------------------------------------------------------- metrics
def stress_normalized(xyz, ti, tj, td):
    return lib().jco_stress_normalized(_f32(xyz), _u32(ti), _u32(tj), _f32(td), len(ti))


def edge_length_cv(xyz, ei, ej):
    return lib().jco_edge_length_cv(_f32(xyz), _u32(ei), _u32(ej), len(ei))


# ---------------------------------------------------------------- multilevel
# crates/graph-layouts/src/layout/coarsen.rs and
# crates/graph-compute/src/engines/multilevel.rs.  The arithmetic (matching,
# fold-up means, prolongation jitter, level CSR) is in jc_oracle.c; the cascade
# loop and the tick state machine are restated here.

def _pairs(edges):
    e = np.asarray(edges, np.uint32).reshape(-1)
    return e if len(e) else np.zeros(2, np.uint32), len(e) // 2


def contract_level(n, edges):
    """coarsen.rs:96-160.  Returns (n_super, parent_map, contracted flat edges)."""
    e, m = _pairs(edges)
    parent = np.zeros(max(n, 1), np.uint32)
    out = np.zeros(max(2 * m, 2), np.uint32)
    cnt = np.zeros(1, np.uint64)
    ns = lib().jco_contract_level(n, e, m, parent, out, cnt)
    return int(ns), parent[:n].copy(), out[: 2 * int(cnt[0])].copy()


def coarsen(n, edges, max_levels, target_size):
    """coarsen.rs:53-86: a list of levels {n, edges, parent_map} (parent_map on
    the coarser level, indexed by the finer level's nodes)."""
    levels = [{"n": n, "edges": np.asarray(edges, np.uint32).reshape(-1), "parent_map": None}]
    for _ in range(1, max_levels):
        cur = levels[-1]
        if cur["n"] <= target_size:
            break
        nn, pm, ne = contract_level(cur["n"], cur["edges"])
        ratio = np.float32(cur["n"]) / np.float32(max(nn, 1))
        if ratio < np.float32(1.6) and nn == cur["n"]:
            break
        levels.append({"n": nn, "edges": ne, "parent_map": pm})
        if nn <= target_size or ratio < np.float32(1.6):
            break
    return levels


def fold_up(levels, xyz):
    """multilevel.rs:486-511: the coarsest level's positions, children averaged."""
    cur = _f32(xyz).copy()
    for lv in levels[1:]:
        out = np.zeros(max(3 * lv["n"], 3), np.float32)
        lib().jco_fold_up_level(len(lv["parent_map"]), cur if len(cur) else np.zeros(3, np.float32),
                                lv["parent_map"] if len(lv["parent_map"]) else np.zeros(1, np.uint32),
                                lv["n"], out)
        cur = out[: 3 * lv["n"]]
    return cur


def prolong(parent_pos, parent_map, n_child, jitter, seed):
    """coarsen.rs:166-193."""
    out = np.zeros(max(3 * n_child, 3), np.float32)
    pm = np.asarray(parent_map, np.uint32)
    lib().jco_prolong(_f32(parent_pos), pm if len(pm) else np.zeros(1, np.uint32), n_child, jitter,
                      seed & 0xFFFFFFFF, out)
    return out[: 3 * n_child]


def level_csr(n, edges):
    """multilevel.rs:229-272."""
    e, m = _pairs(edges)
    off = np.zeros(n + 1, np.uint32)
    nb = np.zeros(max(2 * m, 1), np.uint32)
    lib().jco_level_csr(n, e, m, off, nb)
    return off, nb[: int(off[n])].copy()


def sweeps_for_level(schedule, level, base):
    """multilevel.rs:101-126."""
    base = max(int(base), 1)
    if schedule == "uniform":
        s = base
    elif schedule == "linear":
        s = base * (1 + level)
    else:
        s = base * (1 << min(level, 16))
    return max(min(s, 0xFFFFFFFF), 1)


ML_DEFAULTS = {"max_levels": 6, "target_size": 500, "sweeps_per_level": 30, "sweep_schedule": "linear",
               "spring_len": 30.0, "seed": 0x5EED}


def multilevel_run(offsets, nbrs, xyz, ticks, settings=None, inner="fa2-brute-hip", fa2=None, sgd=None,
                   threads=8):
    """The multilevel wrapper's tick loop (multilevel.rs:333-468) over the
    oracle FA2 (fa2_run with exact sums: the Morton tree for "fa2-bh-hip", no
    tree for "fa2-brute-hip", the default as in multilevel.rs:176) or the
    oracle SGD.  Returns (positions of the current level, level, n_levels)."""
    st = dict(ML_DEFAULTS)
    st.update(settings or {})
    n = len(offsets) - 1
    off = np.asarray(offsets, np.uint32)
    nb = np.asarray(nbrs, np.uint32)
    rows = np.repeat(np.arange(n, dtype=np.uint32), np.diff(off.astype(np.int64)))
    keep = rows < nb
    edges = np.stack([rows[keep], nb[keep]], 1).reshape(-1)
    levels = coarsen(n, edges, st["max_levels"], st["target_size"])
    level = len(levels) - 1
    pos = fold_up(levels, xyz)

    def bind(lv, p):
        lo, lnb = level_csr(levels[lv]["n"], levels[lv]["edges"])
        if inner in ("fa2-bh-hip", "fa2-brute-hip"):
            return {"off": lo, "nb": lnb, "old": None, "se": None}
        k_set = dict(n_pivots=50, eta_max=1.0, eta_min=0.01, n_anneal_steps=30, seed=0x5EEDC0DE12345678)
        k_set.update(sgd or {})
        piv, dist = select_pivots(lo, lnb, k_set["n_pivots"], k_set["seed"])
        em = k_set["eta_max"]
        if em <= 0:
            dm = float(np.float32(dist.max())) if len(dist) else 0.0
            em = float(np.float32(dm) * np.float32(dm))
        return {**k_set, "piv": piv, "dist": dist, "eta_max": em, "t": 0}

    def run(state, p, k):
        if inner in ("fa2-bh-hip", "fa2-brute-hip"):
            p, old, se, _ = fa2_run(state["off"], state["nb"], p, k, fa2 or Fa2Settings.make(),
                                    tree_kind=1 if inner == "fa2-bh-hip" else 2, exact=True,
                                    threads=threads, old_force=state["old"], se=state["se"])
            state["old"], state["se"] = old, se
            return p
        for _ in range(k):
            eta = anneal_eta(state["t"], state["eta_max"], state["eta_min"], state["n_anneal_steps"])
            p = sgd_sweep(state["piv"], state["dist"], p, eta, threads=threads)
            state["t"] += 1
        return p

    state = bind(level, pos)
    done = 0
    left = ticks
    while left > 0:
        if level == 0:
            pos = run(state, pos, left)
            left = 0
            break
        budget = sweeps_for_level(st["sweep_schedule"], level, st["sweeps_per_level"])
        k = min(left, budget - done)
        pos = run(state, pos, k)
        done += k
        left -= k
        if done >= budget:
            child = level - 1
            pos = prolong(pos, levels[level]["parent_map"], levels[child]["n"],
                          np.float32(st["spring_len"]) * np.float32(0.5), st["seed"] + child + 1)
            level = child
            done = 0
            state = bind(level, pos)
    return pos, level, len(levels)

