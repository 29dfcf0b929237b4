"""Pins the CPU oracle before it is trusted as the checker.

Each check is chosen so that a plausible restatement mistake (dropped term,
wrong sign, swapped operand, wrong summation order) fails it: the reference's
own known-answer tests, closed forms, brute force on small inputs, and an
independent exact rational sum.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O


# ---------------------------------------------------------------- SplitMix64
def test_splitmix64_published_sequence():
    # splitmix64.c (Vigna) from state 0: the published first outputs.
    got = O.splitmix_u64(0, 3)
    assert [int(x) for x in got] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_next_f32_closed_form():
    # stability.rs:22-25: (x >> 40) / 2^24
    u = O.splitmix_u64(123, 4)
    xyz = O.uniform_cube(1, 1.0, 123)
    want = [np.float32(2.0) * np.float32(int(v) >> 40) / np.float32(1 << 24) - np.float32(1.0)
            for v in u[:3]]
    assert np.array_equal(xyz, np.array(want, np.float32))


# ---------------------------------------------------------------- speed controller
def test_speed_rise_capped_at_50_percent():
    """fa2_speed.rs:109-123."""
    se = np.array([1.0, 1.0], np.float32)
    stats = np.tile(np.array([0.001, 10.0], np.float32), 100)
    prev = 1.0
    for _ in range(20):
        se, _ = O.speed_update(se, 100, 1.0, stats)
        assert se[0] <= prev * 1.5 + 1e-3
        assert np.isfinite(se[0]) and se[0] > 0
        prev = float(se[0])
    assert se[0] > 1.0


def test_erratic_swing_collapses_speed():
    """fa2_speed.rs:126-138."""
    se = np.array([1.0, 1.0], np.float32)
    stats = np.tile(np.array([25.0, 1.0], np.float32), 100)
    for _ in range(30):
        se, _ = O.speed_update(se, 100, 1.0, stats)
    assert se[0] < 0.05


def test_degenerate_stats_keep_speed():
    """fa2_speed.rs:140-147."""
    for stats, n in ((np.zeros(20, np.float32), 10), (np.full(20, np.nan, np.float32), 10),
                     (np.zeros(0, np.float32), 0)):
        se, _ = O.speed_update(np.array([1.0, 1.0], np.float32), n, 1.0, stats)
        assert se[0] == 1.0 and se[1] == 1.0


def test_speed_first_step_hand_computed():
    # one step from (1, 1) with n=4, swing sum 8, traction sum 2 (erratic branch):
    # est=0.05*2=0.1, min_jt=sqrt(0.1), jt=max(min_jt, min(10, 0.1*2/16)) = sqrt(0.1);
    # ratio 4 > 2 -> eff 0.5, jt=max(jt,1)=1; target = 1*0.5*2/8 = 0.125;
    # swing 8 > 1*2 -> eff 0.35; speed = 1 + min(0.125-1, 0.5) = 0.125
    stats = np.array([2, 0.5] * 4, np.float32)
    se, sums = O.speed_update(np.array([1.0, 1.0], np.float32), 4, 1.0, stats)
    assert sums.tolist() == [8.0, 2.0]
    assert se[0] == np.float32(0.125)
    assert se[1] == np.float32(np.float32(0.5) * np.float32(0.7))


# ---------------------------------------------------------------- exact sums
def test_exact_sum_matches_rational_sum():
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.random(3000).astype(np.float32) * np.float32(1e6),
        (rng.random(500) * 1e-38).astype(np.float32),        # subnormals
        np.array([3.4e38, 1.0, 2.0 ** -149, 0.0, np.nan, np.inf], np.float32),
        rng.random(2000).astype(np.float32),
    ])
    stats = np.stack([vals, vals[::-1]], axis=1).reshape(-1).astype(np.float32)
    got = O.exact_sum2(stats)
    pairs = stats.reshape(-1, 2)
    ok = np.isfinite(pairs).all(axis=1)
    for k in range(2):
        s = sum((Fraction(float(v)) for v in pairs[ok, k]), Fraction(0))
        assert got[k] == float(s)  # float(Fraction) rounds to nearest, ties to even


def test_exact_sum_differs_from_sequential_by_rounding_only():
    rng = np.random.default_rng(1)
    stats = (rng.random(200000) * 10).astype(np.float32)
    ex = O.exact_sum2(stats)
    seq = stats.reshape(-1, 2).astype(np.float64).sum(axis=0)
    assert np.allclose(ex, seq, rtol=1e-12, atol=0)


# ---------------------------------------------------------------- octrees
def _spiral(n):
    ga = np.float32(math.pi * (3.0 - math.sqrt(5.0)))
    out = []
    for i in range(n):
        r = np.float32(math.sqrt(np.float32(i + 1)))
        a = np.float32(i) * ga
        out += [r * np.float32(math.cos(a)), r * np.float32(math.sin(a)), 0.0]
    return np.array(out, np.float32)


def test_reference_octree_theta0_equals_brute_force():
    """fa2_bh.rs:1149-1188: spiral n=16, masses (i%3)+1, scaling 2: leaf mass =
    sum of masses and the theta=0 walk equals brute-force repulsion (1e-3, floor 1)."""
    n = 16
    xyz = _spiral(n)
    mass = np.array([(i % 3) + 1 for i in range(n)], np.float32)
    t = O.octree_ref(xyz, mass)
    leaf = t.body[: t.N] != O.INTERNAL
    assert abs(float(t.node_a[leaf, 3].sum()) - float(mass.sum())) < 1e-3
    # repulsion only: a graph whose 'degree' reproduces the masses, no edges, no gravity
    deg = (mass - 1).astype(np.uint32)
    aoff = np.zeros(n + 1, np.uint32)
    aadj = np.zeros(1, np.uint32)
    s = O.Fa2Settings.make(theta=0.0, gravity=0.0, scaling_ratio=2.0)
    f, _ = O.fa2_forces(xyz, None, deg, aoff, aadj, t, s)
    bf = O.fa2_brute_f64(xyz, deg, aoff, aadj, s)
    rel = np.abs(f.astype(np.float64) - bf) / np.maximum(np.abs(bf), 1.0)
    assert rel.max() < 1e-3


@pytest.mark.parametrize("n,half,seed", [(1, 1.0, 1), (2, 1.0, 2), (3, 5.0, 3), (10, 1.0, 4),
                                         (257, 0.5, 5), (5000, 10.0, 6), (30000, 100.0, 7)])
def test_morton_tree_equals_reference_insertion_tree(n, half, seed):
    xyz = O.uniform_cube(n, half, seed)
    mass = (np.arange(n) % 7 + 1).astype(np.float32)
    a, b = O.octree_ref(xyz, mass, max_nodes=30 * n + 8), O.octree_morton(xyz, mass)
    assert a.N == b.N
    N = a.N
    for f in ("half", "body", "skip", "depth"):
        assert np.array_equal(getattr(a, f)[:N].view(np.uint32), getattr(b, f)[:N].view(np.uint32)), f
    assert np.array_equal(a.node_a.view(np.uint32), b.node_a.view(np.uint32))
    assert np.array_equal(a.root4, b.root4)


def test_morton_tree_equals_reference_on_clustered_layout():
    # positions after 30 FA2 iterations: dense clusters, deep chains
    off, nb = O.ba_graph(3000, 3, 9)
    xyz, _, _, _ = O.fa2_run(off, nb, O.uniform_cube(3000, 1.0, 10), 30, O.Fa2Settings.make(theta=1.0))
    mass = np.ones(3000, np.float32)
    a, b = O.octree_ref(xyz, mass, max_nodes=30 * 3000), O.octree_morton(xyz, mass)
    assert a.N == b.N
    assert np.array_equal(a.node_a.view(np.uint32), b.node_a.view(np.uint32))
    assert np.array_equal(a.skip[: a.N], b.skip[: b.N])


def test_morton_keys_sorted_and_preorder_invariants():
    xyz = O.uniform_cube(4000, 3.0, 12)
    mass = np.ones(4000, np.float32)
    t = O.octree_morton(xyz, mass)
    assert np.all(np.diff(t.keys) >= 0)
    assert np.array_equal(np.sort(t.perm), np.arange(4000, dtype=np.uint32))
    N = t.N
    leaf = t.body[:N] != O.INTERNAL
    assert leaf.sum() == 4000
    # leaves appear in sorted-key order
    assert np.array_equal(t.body[:N][leaf], t.perm)
    # skip of a node is strictly after it; subtree [k, skip) nests
    sk = t.skip[:N].astype(np.int64)
    sk[sk == O.END] = N
    assert np.all(sk > np.arange(N))
    # root mass = total mass
    assert t.node_a[0, 3] == np.float32(4000.0)


# ---------------------------------------------------------------- forces
def test_two_body_closed_form():
    # no edges, no gravity: |F| = k_r (d1+1)(d2+1) / r, directed away
    xyz = np.array([0, 0, 0, 3, 4, 0], np.float32)
    deg = np.array([2, 5], np.uint32)
    aoff = np.zeros(3, np.uint32)
    t = O.octree_morton(xyz, (deg + 1).astype(np.float32))
    s = O.Fa2Settings.make(theta=0.8, gravity=0.0, scaling_ratio=2.0)
    f, _ = O.fa2_forces(xyz, None, deg, aoff, np.zeros(1, np.uint32), t, s)
    f = f.reshape(2, 3).astype(np.float64)
    want = 2.0 * 3 * 6 / 5.0
    assert abs(np.linalg.norm(f[0]) - want) < 1e-5 * want
    assert np.allclose(f[0] / np.linalg.norm(f[0]), [-0.6, -0.8, 0], atol=1e-6)
    assert np.allclose(f[0], -f[1], rtol=1e-6)


def test_gravity_and_attraction_closed_form():
    # isolated body: F = -p/|p| * g * (deg+1) ; edge pair: attraction = d (weight 1)
    xyz = np.array([3, 0, 4, 0, 0, 0], np.float32)
    off, nb = O.csr_from_pairs(2, np.zeros(0, np.uint32))
    deg, aoff, aadj = O.fa2_prepare(2, off, nb)
    t = O.octree_morton(xyz, np.ones(2, np.float32))
    s = O.Fa2Settings.make(scaling_ratio=0.0, gravity=2.0)
    f, _ = O.fa2_forces(xyz, None, deg, aoff, aadj, t, s)
    assert np.allclose(f[:3], [-1.2, 0, -1.6], atol=1e-6)
    off, nb = O.csr_from_pairs(2, np.array([0, 1], np.uint32))
    deg, aoff, aadj = O.fa2_prepare(2, off, nb)
    s = O.Fa2Settings.make(scaling_ratio=0.0, gravity=0.0)
    f, _ = O.fa2_forces(xyz, None, deg, aoff, aadj, t, s)
    assert np.array_equal(f, np.array([-3, 0, -4, 3, 0, 4], np.float32))


def test_attraction_order_is_reference_edge_scan_order():
    # non-symmetric, unsorted rows: the adjacency follows the edge list built row
    # by row from (v, u) with v < u (fa2_bh.rs:264-287), not CSR row order
    off = np.array([0, 2, 4, 5, 6], np.uint32)
    nb = np.array([3, 1, 0, 2, 1, 0], np.uint32)  # row0: 3,1  row1: 0,2  row2: 1  row3: 0
    deg, aoff, aadj = O.fa2_prepare(4, off, nb)
    # edge list: (0,3),(0,1),(1,2) ; row 2 entry (2,1) and row 3 (3,0) have v>u
    assert deg.tolist() == [2, 2, 1, 1]
    assert aoff.tolist() == [0, 2, 4, 5, 6]
    assert aadj.tolist() == [3, 1, 0, 2, 1, 0]


def test_fp64_walk_agrees_with_f32_walk():
    off, nb = O.ba_graph(2000, 3, 21)
    xyz = O.uniform_cube(2000, 1.0, 22)
    deg, aoff, aadj = O.fa2_prepare(2000, off, nb)
    t = O.octree_morton(xyz, deg.astype(np.float32) + 1)
    s = O.Fa2Settings.make(theta=1.0)
    f32, _ = O.fa2_forces(xyz, None, deg, aoff, aadj, t, s)
    f64, flips = O.fa2_forces_f64(xyz, deg, aoff, aadj, t, s)
    err = np.abs(f32 - f64).max() / np.abs(f64).max()
    assert err < 1e-5
    assert flips < 2000 * 5


# ---------------------------------------------------------------- whole runs
def test_reference_path_and_morton_path_runs_agree():
    off, nb = O.sbm_graph(600, 3, 0.05, 0.002, 3)
    xyz = O.uniform_cube(600, 1.0, 4)
    s = O.Fa2Settings.make(theta=1.2)
    a, _, sa, _ = O.fa2_run(off, nb, xyz, 60, s, tree_kind=0, exact=False)
    b, _, sb, _ = O.fa2_run(off, nb, xyz, 60, s, tree_kind=1, exact=True)
    assert np.isfinite(a).all()
    # same tree; only the f64 summation order of the speed sums differs
    assert np.abs(a - b).max() <= 1e-3 * np.abs(a).max()


# ---------------------------------------------------------------- PageRank
def _star(L):
    pairs = np.array([[0, i] for i in range(1, L + 1)], np.uint32).reshape(-1)
    return O.csr_from_pairs(L + 1, pairs)


def test_pagerank_star_closed_form():
    """gpu_pagerank_proptest.rs:70-93: r_c = (1+dL)/(n(1+d)), r_l = (1-d)/n + (d/L) r_c."""
    L, d = 12, 0.85
    off, nb = _star(L)
    r, it, res = O.pagerank_pull(off, nb, d, 200, 1e-10)
    n = L + 1
    rc = (1 + d * L) / (n * (1 + d))
    rl = (1 - d) / n + d / L * rc
    assert abs(r[0] - rc) < 1e-6
    assert np.abs(r[1:] - rl).max() < 1e-6


def test_pagerank_ring_uniform_and_push_equals_pull():
    n = 64
    pairs = np.array([[i, (i + 1) % n] for i in range(n)], np.uint32).reshape(-1)
    off, nb = O.csr_from_pairs(n, pairs)
    r, _, _ = O.pagerank_pull(off, nb, 0.85, 50, 0.0)
    assert np.abs(r - 1.0 / n).max() < 1e-7
    off2, nb2 = O.ba_graph(500, 3, 1)
    a, _, _ = O.pagerank_pull(off2, nb2, 0.85, 50, 0.0)
    b = O.pagerank_push(off2, nb2, 0.85, 50)
    c = O.pagerank_f64(off2, nb2, 0.85, 50)
    assert np.abs(a - b).max() < 1e-6
    assert np.abs(a - c).max() < 1e-6
    assert abs(float(a.sum()) - 1.0) < 1e-4


def test_pagerank_cross_oracle_top20():
    """gpu_pagerank_cross_oracle.rs:89-146 shape: ring + chords + hub, n=300."""
    n = 300
    pairs = [[i, (i + 1) % n] for i in range(n)] + [[i, (i * 7 + 3) % n] for i in range(0, n, 5)]
    pairs += [[0, i] for i in range(2, n, 3)]
    off, nb = O.csr_from_pairs(n, np.array(pairs, np.uint32).reshape(-1))
    a, _, _ = O.pagerank_pull(off, nb, 0.85, 50, 0.0)
    c = O.pagerank_f64(off, nb, 0.85, 50)
    assert np.abs(a - c).max() < 1e-3
    assert np.array_equal(np.argsort(-a, kind="stable")[:20], np.argsort(-c, kind="stable")[:20])


def test_pagerank_rejects_dangling():
    off = np.array([0, 1, 1], np.uint32)
    with pytest.raises(ValueError):
        O.pagerank_pull(off, np.array([1], np.uint32))


# ---------------------------------------------------------------- SGD
def _path(n):
    return O.csr_from_pairs(n, np.array([[i, i + 1] for i in range(n - 1)], np.uint32).reshape(-1))


def test_bfs_path_distances():
    """sgd_stress.rs:411-416."""
    off, nb = _path(5)
    assert O.bfs(off, nb, 0).tolist() == [0, 1, 2, 3, 4]


def test_pivots_are_max_min_brute_force():
    off, nb = O.ba_graph(300, 2, 13)
    piv, dist = O.select_pivots(off, nb, 12, seed=77)
    n = 300
    d = dist.reshape(n, 12).astype(np.int64)
    # recompute max-min greedily with full BFS rows
    rows = [O.bfs(off, nb, int(p)).astype(np.int64) for p in piv]
    mind = np.full(n, np.iinfo(np.int64).max)
    for t in range(11):
        mind = np.minimum(mind, rows[t])
        cand = np.nonzero(mind == mind.max())[0][0]
        assert piv[t + 1] == cand
    for t in range(12):
        assert np.array_equal(np.where(rows[t] > 0, rows[t], 0), d[:, t])
    first = int(O.splitmix_u64(77, 1)[0]) % n
    assert piv[0] == first


def test_sgd_path_is_fixed_point():
    """solved_layouts.rs:254-306: a path laid out on a line at unit spacing has zero stress."""
    n = 24
    off, nb = _path(n)
    piv, dist = O.select_pivots(off, nb, 8)
    pos = np.zeros(3 * n, np.float32)
    pos[0::3] = np.arange(n, dtype=np.float32)
    out = O.sgd_sweep(piv, dist, pos, 1.0)
    assert np.abs(out - pos).max() < 1e-5


def test_anneal_eta():
    """sgd_stress.rs:324-334."""
    assert O.anneal_eta(0, 100.0, 0.01, 30) == np.float32(100.0)
    assert O.anneal_eta(30, 100.0, 0.01, 30) == np.float32(0.01)
    assert O.anneal_eta(5, 1.0, 0.01, 0) == np.float32(0.01)
    v = O.anneal_eta(15, 100.0, 0.01, 30)
    assert abs(v - 1.0) < 1e-4


# ---------------------------------------------------------------- metrics
def test_stress_collinear_path_zero_and_scale_invariant():
    """metrics.rs:205-282."""
    n = 10
    pos = np.zeros(3 * n, np.float32)
    pos[0::3] = np.arange(n, dtype=np.float32) * 3.0
    ti, tj, td = [], [], []
    for i in range(n):
        for j in range(i + 1, n):
            ti.append(i); tj.append(j); td.append(j - i)
    ti, tj, td = (np.array(ti, np.uint32), np.array(tj, np.uint32), np.array(td, np.float32))
    assert O.stress_normalized(pos, ti, tj, td) < 1e-12
    rng = np.random.default_rng(3)
    p2 = rng.random(3 * n).astype(np.float32)
    a = O.stress_normalized(p2, ti, tj, td)
    b = O.stress_normalized((p2 * 7).astype(np.float32), ti, tj, td)
    assert abs(a - b) < 1e-5 and a > 0


def test_edge_length_cv_hand_computed():
    pos = np.array([0, 0, 0, 1, 0, 0, 1, 3, 0], np.float32)
    ei = np.array([0, 1], np.uint32)
    ej = np.array([1, 2], np.uint32)
    # lengths 1 and 3: mean 2, std 1 -> 0.5
    assert abs(O.edge_length_cv(pos, ei, ej) - 0.5) < 1e-7


# ---------------------------------------------------------------- fixtures
def test_csr_from_pairs_semantics():
    pairs = np.array([0, 1, 1, 0, 2, 2, 3, 1, 9, 0, 1, 3], np.uint32)
    off, nb = O.csr_from_pairs(4, pairs)
    assert off.tolist() == [0, 1, 3, 3, 4]
    assert nb.tolist() == [1, 0, 3, 1]


def test_ba_generator_invariants():
    n, m = 2000, 4
    pairs = O.ba_pairs(n, m, 5).reshape(-1, 2)
    assert tuple(pairs[0]) == (0, 1)
    # node v attaches min(m, v) distinct earlier nodes (<= 64 draws can cut it short)
    for v in range(2, n):
        ts = pairs[pairs[:, 0] == v, 1]
        assert len(ts) <= min(m, v) and len(set(ts.tolist())) == len(ts)
        assert np.all(ts < v)
    off, nb = O.csr_from_pairs(n, pairs.reshape(-1))
    for v in range(n):
        row = nb[off[v]:off[v + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0) and v not in row


def test_sbm_generator_against_independent_sampler():
    n, B, pin, pout, seed = 90, 3, 0.2, 0.03, 17
    got = O.sbm_pairs(n, B, pin, pout, seed).reshape(-1, 2).tolist()
    st = [seed]

    def nxt():
        st[0] = (st[0] + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
        z = st[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        return z ^ (z >> 31)

    want = []
    start = [(b * n + B - 1) // B for b in range(B + 1)]
    for a in range(B):
        for b in range(a, B):
            p = pin if a == b else pout
            cand = ([(i, j) for i in range(start[a], start[a + 1]) for j in range(i + 1, start[a + 1])]
                    if a == b else
                    [(i, j) for i in range(start[a], start[a + 1]) for j in range(start[b], start[b + 1])])
            idx = -1
            lq = math.log1p(-p)
            while True:
                u = (nxt() >> 11) * (1.0 / 9007199254740992.0)
                idx += 1 + int(math.floor(math.log1p(-u) / lq))
                if idx >= len(cand):
                    break
                want.append(list(cand[idx]))
    assert got == want


def test_rgg_generator_brute_force():
    n, r = 400, 0.08
    xy, pairs = O.rgg_pairs(n, r, 5)
    p = xy.reshape(-1, 2).astype(np.float64)
    d2 = ((p[:, None, :] - p[None, :, :]) ** 2).sum(-1)
    want = {(i, j) for i in range(n) for j in range(i + 1, n) if d2[i, j] < r * r}
    got = {tuple(x) for x in pairs.reshape(-1, 2).tolist()}
    assert got == want
