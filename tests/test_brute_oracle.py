"""Pins of the oracle's f32 brute-force pass, the fa2-brute engine's fa2_force
(force_atlas2.wgsl:69-135), and of its whole iteration (fa2_brute.rs:454-565).

What fixes it, independently of its own code:
  * the theta=0 tree walk (pinned against the reference's golden behaviour in
    test_oracle.py) visits exactly the leaves, in tree pre-order: where that
    order is the index order the two f32 passes are bit-identical, and where it
    is the reverse they differ in the last bits -- so the brute pass sums in
    index order, with the walk's term;
  * the FP64 brute force, closed forms for two bodies and for one, and the
    reference's own solved-case canaries for fa2-brute (gpu_engines.rs:177-224).
"""
import numpy as np
import pytest

from jumpcannon import workloads
from oracle import oracle as O


def _diag(n, step, reverse=False):
    c = (np.arange(n, dtype=np.float64) * step - 0.5 * n * step).astype(np.float32)
    if reverse:
        c = c[::-1].copy()
    return np.repeat(c, 3).astype(np.float32)


def _prep(n, off, nb):
    return O.fa2_prepare(n, off, nb)


@pytest.mark.parametrize("kw", [{}, {"lin_log_mode": 1, "strong_gravity": 1},
                                {"scaling_ratio": 0.37, "gravity": 0.0}])
def test_brute_equals_theta0_walk_when_tree_order_is_index_order(kw):
    n = 400
    off, nb = O.ba_graph(n, 3, 41)
    deg, aoff, aadj = _prep(n, off, nb)
    xyz = _diag(n, 0.05)
    t = O.octree_morton(xyz, deg.astype(np.float32) + 1)
    assert np.array_equal(t.perm, np.arange(n, dtype=np.uint32))  # the test's premise
    s = O.Fa2Settings.make(theta=0.0, **kw)
    old = O.uniform_cube(n, 1.0, 42)
    fw, sw = O.fa2_forces(xyz, old, deg, aoff, aadj, t, s)
    fb, sb = O.fa2_brute_forces(xyz, old, deg, aoff, aadj, s)
    assert np.array_equal(fw.view(np.uint32), fb.view(np.uint32))
    assert np.array_equal(sw.view(np.uint32), sb.view(np.uint32))


def test_brute_sums_in_index_order_not_tree_order():
    n = 400
    off, nb = O.ba_graph(n, 3, 43)
    deg, aoff, aadj = _prep(n, off, nb)
    xyz = _diag(n, 0.05, reverse=True)
    t = O.octree_morton(xyz, deg.astype(np.float32) + 1)
    assert np.array_equal(t.perm, np.arange(n, dtype=np.uint32)[::-1])
    s = O.Fa2Settings.make(theta=0.0, gravity=0.0)
    aoff0 = np.zeros(n + 1, np.uint32)  # repulsion only
    fw, _ = O.fa2_forces(xyz, None, deg, aoff0, aadj, t, s)
    fb, _ = O.fa2_brute_forces(xyz, None, deg, aoff0, aadj, s)
    assert not np.array_equal(fw, fb)
    assert np.abs(fw - fb).max() <= 1e-5 * np.abs(fb).max()


def test_brute_matches_fp64_brute_force():
    n = 700
    off, nb = O.ba_graph(n, 3, 44)
    deg, aoff, aadj = _prep(n, off, nb)
    xyz = O.uniform_cube(n, 2.0, 45)
    for kw in ({}, {"lin_log_mode": 1}, {"strong_gravity": 1, "scaling_ratio": 5.0}):
        s = O.Fa2Settings.make(**kw)
        f32, _ = O.fa2_brute_forces(xyz, None, deg, aoff, aadj, s)
        f64 = O.fa2_brute_f64(xyz, deg, aoff, aadj, s)
        assert np.abs(f32 - f64).max() / np.abs(f64).max() < 1e-5, kw


def test_brute_two_body_and_one_body_closed_forms():
    # |F| = k_r (d1+1)(d2+1) / r along d = p_i - p_j (force_atlas2.wgsl:82-91)
    xyz = np.array([0, 0, 0, 3, 4, 0], np.float32)
    deg = np.array([2, 5], np.uint32)
    aoff = np.zeros(3, np.uint32)
    s = O.Fa2Settings.make(gravity=0.0, scaling_ratio=2.0)
    f, st = O.fa2_brute_forces(xyz, None, deg, aoff, np.zeros(1, np.uint32), s)
    f = f.reshape(2, 3).astype(np.float64)
    want = 2.0 * 3 * 6 / 5.0
    assert abs(np.linalg.norm(f[0]) - want) < 1e-5 * want
    assert np.allclose(f[0] / np.linalg.norm(f[0]), [-0.6, -0.8, 0], atol=1e-6)
    assert np.allclose(f[0], -f[1], rtol=1e-6)
    # F_{t-1} = 0: swing = deg*|F|, traction = deg*|F|/2 (force_atlas2.wgsl:129-133)
    assert np.allclose(st[0::2], [3 * want, 6 * want], rtol=1e-6)
    assert np.allclose(st[1::2], [1.5 * want, 3 * want], rtol=1e-6)
    # closer than 0.01: |d|^2 clamps to 1e-4, so |F| = k_r (d1+1)(d2+1) |d| / 1e-4
    near = np.array([0, 0, 0, 0.003, 0.004, 0], np.float32)
    f, _ = O.fa2_brute_forces(near, None, deg, aoff, np.zeros(1, np.uint32), s)
    want = 2.0 * 3 * 6 * 0.005 / 1e-4
    assert abs(np.linalg.norm(f[:3].astype(np.float64)) - want) < 1e-4 * want
    # one body: no repulsion (j == i is skipped), gravity -p/|p| * g * (deg+1)
    f, _ = O.fa2_brute_forces(np.array([3, 0, 4], np.float32), None, np.zeros(1, np.uint32),
                              np.zeros(2, np.uint32), np.zeros(1, np.uint32),
                              O.Fa2Settings.make(gravity=2.0))
    assert np.allclose(f, [-1.2, 0, -1.6], atol=1e-6)


def test_non_finite_body_reaches_every_other_body():
    """The brute pass has no finiteness filter (force_atlas2.wgsl:82-91): one
    NaN body makes every other body's repulsion NaN; the NaN body itself skips
    only its own term."""
    n = 50
    off, nb = O.ba_graph(n, 2, 46)
    deg, aoff, aadj = _prep(n, off, nb)
    xyz = O.uniform_cube(n, 1.0, 47)
    xyz[3 * 7 + 1] = np.nan
    f, _ = O.fa2_brute_forces(xyz, None, deg, aoff, aadj, O.Fa2Settings.make())
    f = f.reshape(n, 3)
    assert np.isnan(f).any(axis=1).all()


def test_brute_run_is_force_speed_apply():
    """tree_kind 2 of the oracle's run = brute pass, AdaptiveSpeed on the exact
    sums, apply (fa2_brute.rs:454-565), iteration after iteration."""
    off, nb, xyz = workloads.sbm_2000()
    n = len(off) - 1
    deg, aoff, aadj = _prep(n, off, nb)
    s = O.Fa2Settings.make()
    want, old_w, se_w, _ = O.fa2_run(off, nb, xyz, 4, s, tree_kind=2, exact=True)
    x, old, se = xyz.copy(), np.zeros(3 * n, np.float32), np.array([1.0, 1.0], np.float32)
    for _ in range(4):
        f, st = O.fa2_brute_forces(x, old, deg, aoff, aadj, s)
        se, _ = O.speed_update(se, n, s.jitter_tolerance, st, exact=True)
        x, old = O.fa2_apply(x, f, old, deg, s, float(se[0]))
    assert np.array_equal(x, want) and np.array_equal(old, old_w) and np.array_equal(se, se_w)


def _dist(p, i, j):
    return float(np.linalg.norm(p[3 * i:3 * i + 3].astype(np.float64) - p[3 * j:3 * j + 3]))


def test_reference_canaries_on_the_oracle():
    """gpu_engines.rs:177-224 for fa2-brute: a single edge stays finite for 400
    steps; K3 from a scalene seed relaxes to equilateral within 15% in 600."""
    off, nb = O.csr_from_pairs(2, np.array([0, 1], np.uint32))
    x = np.array([0, 0, 0, 5, 0, 0], np.float32)
    old, se = None, None
    for _ in range(400):
        x, old, se, _ = O.fa2_run(off, nb, x, 1, O.Fa2Settings.make(), tree_kind=2, exact=False,
                                  old_force=old, se=se)
        assert np.isfinite(x).all()
    off = np.array([0, 2, 4, 6], np.uint32)
    nb = np.array([1, 2, 0, 2, 0, 1], np.uint32)
    x = np.array([0, 0, 0, 10, 0, 0, 3, 4, 0], np.float32)
    out, _, _, _ = O.fa2_run(off, nb, x, 600, O.Fa2Settings.make(), tree_kind=2, exact=False)
    d = [_dist(out, 0, 1), _dist(out, 1, 2), _dist(out, 2, 0)]
    mean = sum(d) / 3
    assert mean > 1e-3 and max(abs(v - mean) for v in d) / mean < 0.15
