"""Closed-form pins of the oracle pieces that the whole-run comparisons leave
open: the octree's octant digits and cell geometry, the apply step's
displacement, the SGD per-term step, and the stress and edge-length metrics.

Every expected value is written from the rule as the reference states it, not
by calling the oracle, and each case is chosen so that a plausible slip in the
restatement fails it:
  * y and z octant bits exchanged (octant_for / child_center, fa2_bh.rs:841-861)
    -> test_octree_hand_placed_cells, test_morton_key_leading_digits;
  * a child centre at c +- h instead of c +- h/2
    -> test_octree_hand_placed_cells, test_bodies_inside_every_cell_on_their_path;
  * speed / (1 + speed * sqrt(swing)) in apply (fa2_barnes_hut.wgsl:203-226)
    -> test_apply_closed_form;
  * mu = min(1, eta / d) instead of min(1, eta / d^2) (sgd_stress.wgsl:48-58)
    -> test_sgd_single_term_closed_form;
  * pivots of one node applied Jacobi-style instead of in sequence
    -> test_sgd_two_pivots_in_sequence;
  * a missing / sum(w d^2) in scale_normalized_stress (metrics.rs:31-59)
    -> test_stress_hand_computed.
The reference tree's failure semantics (capacity 2n+8, depth cap 32,
fa2_bh.rs:813-819, 863-923) and where the GPU tree departs from them are
pinned at the end.
"""
import math
from fractions import Fraction

import numpy as np

from oracle import oracle as O

f32 = np.float32


# ---------------------------------------------------------------- octree geometry
# One body on each corner of [-1,1]^3 except (+,+,+), plus P and Q in the
# (+,+,+) octant; P and Q share the cells of depths 1-3 and part at depth 3,
# where P is in octant 3 (x, y) and Q in octant 5 (x, z).  Body ids are
# scrambled so that id order, octant order and pre-order all differ.
CORNER_IDS = [4, 0, 6, 2, 8, 1, 3]  # bodies on the corners of octants 0..6
P_ID, Q_ID = 5, 7
P = (1.0, 1.0, 0.8)
Q = (1.0, 0.8, 1.0)


def _hand_layout():
    xyz = np.zeros((9, 3), np.float32)
    for o, b in enumerate(CORNER_IDS):
        xyz[b] = [1.0 if o & 1 else -1.0, 1.0 if o & 2 else -1.0, 1.0 if o & 4 else -1.0]
    xyz[P_ID] = P
    xyz[Q_ID] = Q
    return xyz.reshape(-1)


def _child(c, h, o):
    """fa2_bh.rs:855-861 as stated: half = parent/2, centre = parent +- half,
    x for bit 1, y for bit 2, z for bit 4."""
    hh = f32(h) * f32(0.5)
    return (tuple(f32(c[k]) + (hh if o & (1 << k) else -hh) for k in range(3)), hh)


def _expected_preorder():
    """(body or None, centre, half, depth) in pre-order, by hand."""
    root_c, root_h = (f32(0), f32(0), f32(0)), f32(1.0) * f32(1.05)  # bbox [-1,1]^3, x1.05
    out = [(None, root_c, root_h, 0)]
    for o in range(7):
        c, h = _child(root_c, root_h, o)
        out.append((CORNER_IDS[o], c, h, 1))
    c1, h1 = _child(root_c, root_h, 7)    # centre .525
    c2, h2 = _child(c1, h1, 7)            # centre .7875: P, Q both >= on every axis
    c3, h3 = _child(c2, h2, 7)            # centre .91875: P.z and Q.y fall below
    out += [(None, c1, h1, 1), (None, c2, h2, 2), (None, c3, h3, 3)]
    cp, hp = _child(c3, h3, 3)
    cq, hq = _child(c3, h3, 5)
    out += [(P_ID, cp, hp, 4), (Q_ID, cq, hq, 4)]
    return out


def test_octree_hand_placed_cells():
    xyz = _hand_layout()
    t = O.octree_ref(xyz, np.ones(9, np.float32))
    exp = _expected_preorder()
    assert t.N == len(exp)
    assert np.array_equal(t.root4, np.array([0, 0, 0, f32(1.0) * f32(1.05)], np.float32))
    for q, (b, c, h, d) in enumerate(exp):
        assert t.body[q] == (O.INTERNAL if b is None else b), q
        assert t.half[q] == h, q
        assert tuple(t.centre[q]) == c, (q, t.centre[q], c)
        assert t.depth[q] == d, q
    # the Morton tree (the GPU's contract) visits the bodies in the same order
    m = O.octree_morton(xyz, np.ones(9, np.float32))
    leaves = m.body[: m.N][m.body[: m.N] != O.INTERNAL].tolist()
    assert leaves == [b for b, _, _, _ in exp if b is not None] == CORNER_IDS + [P_ID, Q_ID]
    assert m.N == t.N and np.array_equal(m.half[: m.N], t.half[: t.N])


def _digits(key, k):
    return [int(key >> (60 - 3 * l)) & 7 for l in range(k)]


def test_morton_key_leading_digits():
    """The key is the octant path through the same cell chain: corners have their
    octant as the first digit, P and Q share 7,7,7 and then split 3 / 5."""
    xyz = _hand_layout()
    keys, root4 = O.morton_keys(xyz)
    for o, b in enumerate(CORNER_IDS):
        assert _digits(int(keys[b]), 1) == [o]
    assert _digits(int(keys[P_ID]), 4) == [7, 7, 7, 3]
    assert _digits(int(keys[Q_ID]), 4) == [7, 7, 7, 5]
    assert int(keys[P_ID]) >> 63 == 0


def test_bodies_inside_every_cell_on_their_path():
    """Each cell [c - h, c + h] holds every body of its subtree, at every depth."""
    for xyz in (O.uniform_cube(3000, 5.0, 21), _hand_layout()):
        n = len(xyz) // 3
        t = O.octree_ref(xyz, np.ones(n, np.float32), max_nodes=40 * n + 8)
        p = xyz.reshape(-1, 3)
        N = t.N
        skip = t.skip[:N].astype(np.int64)
        skip[skip == O.END] = N
        for q in range(N):
            lo, hi = t.centre[q] - t.half[q], t.centre[q] + t.half[q]
            sub = t.body[q:skip[q]]
            b = sub[sub != O.INTERNAL]
            assert len(b) >= 1
            inside = (p[b] >= lo) & (p[b] <= hi)
            assert inside.all(), (q, t.depth[q])
            # a child's centre is its parent's +- half the child's half extent
            if t.depth[q] > 0:
                par = q - 1
                while t.depth[par] != t.depth[q] - 1:
                    par -= 1
                h = t.half[q]
                assert np.all((t.centre[q] == t.centre[par] + h) | (t.centre[q] == t.centre[par] - h))
                assert t.half[q] == t.half[par] * f32(0.5)


# ---------------------------------------------------------------- apply
def _apply_one(force, old, deg, speed, **settings):
    s = O.Fa2Settings.make(**settings)
    xyz = np.array([1.0, 2.0, 3.0], np.float32)
    out, new_old = O.fa2_apply(xyz, np.array(force, np.float32), np.array(old, np.float32),
                               np.array([deg], np.uint32), s, speed)
    assert np.array_equal(new_old, np.array(force, np.float32))  # F_{t-1} <- F
    return out.astype(np.float64) - xyz.astype(np.float64)


def test_apply_closed_form():
    """fa2_barnes_hut.wgsl:203-226: swing = (deg+1)|F_{t-1} - F|,
    factor = s / (1 + sqrt(s * swing)), capped at k_smax / |F|, times time_step."""
    F = [3.0, 4.0, 0.0]                                      # |F| = 5
    d = _apply_one(F, [0, 0, 0], 0, 2.0)                     # swing 5
    fac = 2.0 / (1.0 + math.sqrt(2.0 * 5.0))
    assert np.allclose(d, np.array(F) * fac, rtol=2e-6, atol=0)
    d = _apply_one(F, [0, 0, 0], 0, 2.0, time_step=0.5)
    assert np.allclose(d, np.array(F) * fac * 0.5, rtol=2e-6, atol=0)
    # k_smax cap: factor = min(0.4805, 1 / 5): the step is exactly k_smax long
    d = _apply_one(F, [0, 0, 0], 0, 2.0, max_displacement=1.0)
    assert np.allclose(d, [0.6, 0.8, 0.0], rtol=2e-6, atol=1e-7)
    # k_smax <= 0 switches the cap off
    d = _apply_one([300.0, 400.0, 0.0], [300.0, 400.0, 0.0], 0, 2.0, max_displacement=0.0)
    assert np.allclose(d, [600.0, 800.0, 0.0], rtol=2e-6)      # swing 0: factor = s
    # degree weighting of the swing: deg 2 -> (deg+1) = 3, |F_{t-1} - F| = 1
    d = _apply_one([0, 0, 2.0], [0, 0, 1.0], 2, 0.5)
    assert np.allclose(d, [0, 0, 2.0 * 0.5 / (1.0 + math.sqrt(0.5 * 3.0))], rtol=2e-6)


# ---------------------------------------------------------------- SGD
def test_sgd_single_term_closed_form():
    """sgd_stress.wgsl:48-58: after one term |p' - q| = |D| - mu (|D| - d) with
    mu = min(1, eta / d^2); the pivot (its own row all zero) does not move."""
    piv = np.array([0], np.uint32)
    dist = np.array([0, 3], np.uint16)                       # node-major, k = 1
    pos = np.array([0, 0, 0, 3.0, 4.0, 0], np.float32)       # |D| = 5, d = 3
    for eta, mu in ((2.0, 2.0 / 9.0), (0.9, 0.1), (100.0, 1.0)):
        out = O.sgd_sweep(piv, dist, pos, eta, threads=1).reshape(2, 3).astype(np.float64)
        assert np.array_equal(out[0], [0, 0, 0])
        want = 5.0 - mu * (5.0 - 3.0)
        assert abs(np.linalg.norm(out[1]) - want) < 2e-6 * want, (eta, np.linalg.norm(out[1]), want)
        assert np.allclose(out[1] / np.linalg.norm(out[1]), [0.6, 0.8, 0.0], atol=1e-6)


def test_sgd_two_pivots_in_sequence():
    """One node's pivots are applied one after the other (the second term sees
    the first one's move), pivots read from the sweep's input positions."""
    piv = np.array([0, 1], np.uint32)
    dist = np.array([0, 0, 0, 0, 2, 1], np.uint16)           # node 2: d=2 to pivot 0, d=1 to pivot 1
    pos = np.array([0, 0, 0, 3, 4, 0, 0, 4, 0], np.float32)
    out = O.sgd_sweep(piv, dist, pos, 1.0, threads=1).reshape(3, 3).astype(np.float64)
    p = np.array([0.0, 4.0, 0.0])
    for q, d in (([0.0, 0.0, 0.0], 2.0), ([3.0, 4.0, 0.0], 1.0)):
        D = p - q
        mag = np.linalg.norm(D)
        mu = min(1.0, 1.0 / d ** 2)
        p = p - mu * (mag - d) / mag * D
    assert np.allclose(out[2], p, rtol=2e-6, atol=1e-6)
    assert np.allclose(out[:2], pos.reshape(3, 3)[:2])
    assert abs(np.linalg.norm(out[2] - [3, 4, 0]) - 1.0) < 2e-6  # mu = 1 lands at distance d


# ---------------------------------------------------------------- metrics
def test_stress_hand_computed():
    """Line x = 0, 1, 3; terms (0,1,1), (0,2,2), (1,2,1): alpha = 18/29, stress
    (sum w (alpha D - d)^2) / sum w d^2 = (6/29) / 3 = 2/29 exactly."""
    xyz = np.array([0, 0, 0, 1, 0, 0, 3, 0, 0], np.float32)
    ti, tj, td = np.array([0, 0, 1], np.uint32), np.array([1, 2, 2], np.uint32), np.array([1, 2, 1], np.float32)
    terms = [(1, 1), (2, 3), (1, 2)]  # (d, D)
    alpha = sum(Fraction(1, d * d) * d * D for d, D in terms) / sum(Fraction(1, d * d) * D * D for d, D in terms)
    assert alpha == Fraction(18, 29)
    s = sum(Fraction(1, d * d) * (alpha * D - d) ** 2 for d, D in terms) / sum(1 for _ in terms)
    assert s == Fraction(2, 29)
    assert O.stress_normalized(xyz, ti, tj, td) == float(f32(float(s)))
    # a term with d <= 0 is skipped
    assert O.stress_normalized(xyz, np.append(ti, 0), np.append(tj, 2), np.append(td, 0)) == float(f32(2 / 29))


def test_edge_length_cv_hand_computed():
    """metrics.rs:150-164: lengths 1, 3, 2 -> population sd sqrt(2/3), mean 2."""
    xyz = np.array([0, 0, 0, 1, 0, 0, 3, 0, 0], np.float32)
    cv = O.edge_length_cv(xyz, np.array([0, 0, 1], np.uint32), np.array([1, 2, 2], np.uint32))
    assert cv == float(f32(math.sqrt(2.0 / 3.0) / 2.0))
    assert O.edge_length_cv(xyz, np.array([0], np.uint32), np.array([1], np.uint32)) == 0.0


# ---------------------------------------------------------------- tree failure semantics
def test_reference_capacity_overflow_keeps_a_prefix():
    """fa2_bh.rs:813-819, 881-883: with capacity 2n+8 the reference stops at
    the first body that does not fit and keeps the tree built so far.  Close
    pairs (1e-4 apart) need chains of cells, so 2n+8 overflows and most
    bodies are missing; the GPU's Morton tree (no capacity limit below
    n + 21(n/2+1)) holds every body and equals the uncapped insertion tree."""
    rng = np.random.default_rng(3)
    n = 2000
    c = rng.uniform(-10, 10, (n // 2, 3)).astype(np.float32)
    xyz = np.repeat(c, 2, axis=0)
    xyz[1::2] += np.float32(1e-4)
    xyz = xyz.reshape(-1).astype(np.float32)
    m = np.ones(n, np.float32)
    capped = O.octree_ref(xyz, m)  # the reference's 2n+8
    leaves = capped.body[: capped.N][capped.body[: capped.N] != O.INTERNAL]
    k = len(leaves)
    assert capped.N == 2 * n + 8 and k < n // 2
    assert sorted(leaves.tolist()) == list(range(k))  # a prefix in body order
    assert capped.node_a[0, 3] == np.float32(k)
    full = O.octree_ref(xyz, m, max_nodes=40 * n + 8)
    mort = O.octree_morton(xyz, m)
    assert full.N == mort.N > 2 * n + 8
    assert np.array_equal(full.node_a.view(np.uint32), mort.node_a.view(np.uint32))
    assert np.array_equal(full.skip[: full.N], mort.skip[: mort.N])


def test_coincident_bodies_reference_drops_morton_keeps():
    """fa2_bh.rs:872-874, 922: a body that still collides after 32 levels is
    dropped (each retry pushes the twin one level down), so the reference tree
    loses every later copy of a position.  The Morton tree stops at 21 levels
    and keeps all copies as sibling leaves of one depth-21 cell."""
    n = 50
    xyz = O.uniform_cube(n, 1.0, 4).copy()
    for dup in (7, 9):
        xyz[3 * dup:3 * dup + 3] = xyz[9:12]  # copies of body 3
    m = np.ones(n, np.float32)
    ref = O.octree_ref(xyz, m, max_nodes=40 * n + 8)
    rb = ref.body[: ref.N]
    assert sorted(set(range(n)) - set(rb[rb != O.INTERNAL].tolist())) == [7, 9]
    assert ref.node_a[0, 3] == np.float32(n - 2)
    assert ref.depth[np.nonzero(rb == 3)[0][0]] == 33
    mort = O.octree_morton(xyz, m)
    mb = mort.body[: mort.N]
    assert sorted(mb[mb != O.INTERNAL].tolist()) == list(range(n))
    q = [int(np.nonzero(mb == b)[0][0]) for b in (3, 7, 9)]
    assert q == [q[0], q[0] + 1, q[0] + 2] and all(mort.depth[x] == 22 for x in q)
    assert mort.node_a[0, 3] == np.float32(n)
    # the three copies are the only leaves of their depth-21 cell
    assert mort.depth[q[0] - 1] == 21 and mort.skip[q[0] - 1] == (q[2] + 1 if q[2] + 1 < mort.N else O.END)
