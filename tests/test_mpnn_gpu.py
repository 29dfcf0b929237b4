"""The warm-start network's HIP kernels against the float64 oracle: MFMA
forward pass, gradients of the invariant distance loss, training progress and
run-to-run determinism."""
import numpy as np
import pytest

import jumpcannon as jc
from oracle import mpnn_oracle as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _anchors(seed, n, na):
    s, out = seed, []
    mask = (1 << 64) - 1
    while len(out) < min(na, n):
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        c = z % n
        if c not in out:
            out.append(c)
    return out


def _setup(n=400, seed=5):
    off, nb = O.ba_graph(n, 3, seed)
    net = jc.Mpnn(seed=11)
    net.set_graph(off, nb, feat_seed=9, walks=16)
    return net, off, nb


def test_rwpe_ring_return_probabilities():
    n = 64
    off, nb = O.csr_from_pairs(n, np.array([[i, (i + 1) % n] for i in range(n)], np.uint32).reshape(-1))
    net = jc.Mpnn(seed=1)
    net.set_graph(off, nb, feat_seed=3, walks=4096)
    x = net.features()
    # on a cycle: P(back after 1 step) = 0, after 2 steps = 1/2, after 4 = 3/8
    assert np.all(x[:, 2] == 0.0)
    assert abs(x[:, 3].mean() - 0.5) < 0.01 and abs(x[:, 5].mean() - 0.375) < 0.01
    assert np.allclose(x[:, 0], 1.0) and np.allclose(x[:, 1], np.float32(np.log(3.0)), atol=1e-2)


def test_forward_matches_oracle():
    net, off, nb = _setup()
    y, ms = net.forward()
    x = net.features().astype(np.float64)
    yo, cache = M.forward(net.params(), off, nb, x)
    h0 = net.activations(0)
    agree = np.mean(h0 == cache[0][1].astype(np.float32))
    rel = np.abs(y.reshape(-1, 3) - yo).max() / np.sqrt((yo ** 2).mean())
    print(f"layer-0 bf16 agreement {agree:.4f}, output max rel {rel:.2e}")
    assert agree > 0.98 and rel < 2e-2


def test_gradient_matches_oracle():
    net, off, nb = _setup(300, 6)
    n = len(off) - 1
    tgt = O.uniform_cube(n, 3.0, 12)
    net.set_target(tgt, n_anchor=16, seed=4)
    L = net.loss_grad()
    g = net.grads().astype(np.float64)
    Lo, go = M.loss_grad(net.params(), off, nb, net.features().astype(np.float64), tgt,
                         _anchors(4, n, 16))
    assert abs(L - Lo) <= 1e-3 * abs(Lo)
    o = 0
    for k in list(M.K) + [None]:
        size = (k + 1) * M.HID if k is not None else (M.HID + 1) * 3
        a, b = g[o:o + size], go[o:o + size]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        print(f"block K={k}: rel {rel:.2e} cos {cos:.6f}")
        assert rel < 3e-2 and cos > 0.999
        o += size


def test_training_reduces_loss_and_is_deterministic():
    runs = []
    for _ in range(2):
        net, off, nb = _setup(2000, 8)
        n = len(off) - 1
        xyz = O.uniform_cube(n, 1.0, 2)
        tgt, _, _, _ = O.fa2_run(off, nb, xyz, 150, O.Fa2Settings.make(theta=1.0))
        net.set_target(tgt, n_anchor=32, seed=5)
        L0 = net.loss_grad()
        L1 = net.train(300, 2e-3)
        runs.append((L0, L1, net.params()))
    L0, L1, p = runs[0]
    print(f"loss {L0:.4f} -> {L1:.4f}")
    assert L1 < 0.5 * L0
    assert runs[1][1] == L1 and np.array_equal(runs[1][2], p)


def test_param_access_is_ordered_after_queued_updates():
    """params()/set_params()/grads() right after queued work read or write the
    values that work leaves (copies on the network's stream)."""
    net, off, nb = _setup(2000, 8)
    n = len(off) - 1
    net.set_target(O.uniform_cube(n, 1.0, 9), n_anchor=32, seed=5)
    net.loss_grad()
    g = net.grads()
    net.adam(1e-3)
    p1 = net.params()
    assert np.array_equal(p1, net.params())
    assert np.any(p1 != 0) and np.any(g != 0)
    net.loss_grad()
    net.adam(1e-3)
    net.set_params(p1)  # lands after the queued step, which it overwrites
    assert np.array_equal(net.params(), p1)
