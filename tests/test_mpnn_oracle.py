"""Pins the float64 warm-start network oracle (CPU only): its gradient against
central finite differences, the loss invariances the design relies on, and the
aggregation transpose."""
import numpy as np

from oracle import mpnn_oracle as M
from oracle import oracle as O


def _graph(n=40, m=2, seed=3):
    return O.ba_graph(n, m, seed)


def _params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    P = sum((k + 1) * M.HID for k in M.K) + (M.HID + 1) * 3
    return (rng.standard_normal(P) * 0.1 * scale).astype(np.float64)


def test_aggregation_transpose_identity():
    off, nb = _graph(60)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((60, 5))
    g = rng.standard_normal((60, 5))
    assert abs((M.mean_agg(off, nb, x) * g).sum() - (x * M.mean_agg_t(off, nb, g)).sum()) < 1e-10


def test_gradient_matches_finite_differences():
    off, nb = _graph(30)
    n = 30
    rng = np.random.default_rng(2)
    x = rng.standard_normal((n, 16)) * 0.5
    tgt = rng.standard_normal(3 * n)
    anchors = [0, 7]
    p = _params(4, scale=3.0)
    L, g = M.loss_grad(p, off, nb, x, tgt, anchors, emulate_bf16=False)
    idx = rng.choice(len(p), 40, replace=False)
    h = 1e-6
    for j in idx:
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        Lp, _ = M.loss_grad(pp, off, nb, x, tgt, anchors, emulate_bf16=False)
        Lm, _ = M.loss_grad(pm, off, nb, x, tgt, anchors, emulate_bf16=False)
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 + 1e-4 * abs(fd), (j, fd, g[j])


def test_loss_invariant_to_rigid_motion_and_scale():
    off, nb = _graph(50)
    rng = np.random.default_rng(5)
    y = rng.standard_normal((50, 3))
    t = rng.standard_normal(150)
    pi, pj = M.pairs(off, nb, [3, 11])
    L0, _ = M.loss_and_dy(y, t, pi, pj)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    L1, _ = M.loss_and_dy(3.7 * y @ q + 2.0, t, pi, pj)
    L2, _ = M.loss_and_dy(y, 0.25 * t, pi, pj)
    assert abs(L0 - L1) < 1e-10 and abs(L0 - L2) < 1e-10
    L3, _ = M.loss_and_dy(t.reshape(-1, 3) * 5.0, t, pi, pj)
    assert L3 < 1e-20  # perfect up to scale


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.00390625, 1.005859375, 1.0 + 2 ** -8 + 2 ** -9], np.float32)
    # 1 + 2^-8 is a tie between 1 and 1+2^-7 -> even (1.0)
    assert M.bf16(v).tolist() == [1.0, 1.0, 1.0078125, 1.0078125]
