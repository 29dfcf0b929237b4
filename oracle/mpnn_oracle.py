"""Float64 numpy restatement of the warm-start network (TEST INFRASTRUCTURE ONLY).

Mirrors jump-cannon_amd/csrc/mpnn.hip: the same layers, loss and backward pass,
computed in float64, with BF16 rounding (round to nearest even) applied at the
points where the GPU stores BF16: input features, aggregated means, layer
outputs, the weights fed to MFMA, and the upstream gradient fed to the dX
product.  The network has no reference counterpart; this oracle pins the GPU
forward pass and its gradients (checked against central finite differences of
this same function on small graphs in tests/test_mpnn_oracle.py).
"""
from __future__ import annotations

import numpy as np

HID = 128
IN = 16
K = (2 * IN, 2 * HID, 2 * HID)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even) and return them as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, np.float64)
    out, o = [], 0
    for k in K:
        W = p[o:o + k * HID].reshape(k, HID)
        o += k * HID
        b = p[o:o + HID]
        o += HID
        out.append((W, b))
    Wh = p[o:o + HID * 3].reshape(HID, 3)
    o += HID * 3
    bh = p[o:o + 3]
    return out, Wh, bh


def mean_agg(off, nb, H):
    n = len(off) - 1
    deg = np.diff(off).astype(np.float64)
    src = np.repeat(np.arange(n), np.diff(off))
    S = np.zeros_like(H)
    np.add.at(S, src, H[nb])
    inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    return S * inv[:, None]


def mean_agg_t(off, nb, G):
    """Transpose of mean_agg on a symmetric CSR: out_j = sum_{i in N(j)} G_i / deg_i."""
    n = len(off) - 1
    deg = np.diff(off).astype(np.float64)
    src = np.repeat(np.arange(n), np.diff(off))
    out = np.zeros_like(G)
    np.add.at(out, src, G[nb] / deg[nb][:, None])
    return out


def forward(params, off, nb, x, emulate_bf16=True):
    """x: (n,16) input features as the GPU holds them.  Returns (y, cache)."""
    r = bf16 if emulate_bf16 else (lambda v: np.asarray(v, np.float64))
    layers, Wh, bh = unpack(params)
    H = r(x)
    cache = []
    for W, b in layers:
        A = r(mean_agg(off, nb, H))
        X = np.concatenate([H, A], axis=1)
        Z = X @ r(W) + b
        Hn = r(np.maximum(Z, 0.0))
        cache.append((X, Hn, W))
        H = Hn
    y = H @ Wh + bh
    return y, cache


def pairs(off, nb, anchors):
    n = len(off) - 1
    src = np.repeat(np.arange(n), np.diff(off))
    m = nb > src
    pi = [src[m]]
    pj = [nb[m]]
    for a in anchors:
        i = np.arange(n)
        i = i[i != a]
        pi.append(i)
        pj.append(np.full(len(i), a))
    return np.concatenate(pi), np.concatenate(pj)


def loss_and_dy(y, tgt, pi, pj):
    t = np.asarray(tgt, np.float64).reshape(-1, 3)
    D = np.linalg.norm(t[pi] - t[pj], axis=1)
    dv = y[pi] - y[pj]
    d = np.linalg.norm(dv, axis=1)
    ok = D > 0
    w = np.where(ok, 1.0 / np.where(ok, D, 1.0) ** 2, 0.0)
    num, den, wd2 = (w * D * d).sum(), (w * d * d).sum(), (w * D * D).sum()
    alpha = num / den
    L = (w * (alpha * d - D) ** 2).sum() / wd2
    c = np.where((w > 0) & (d > 0), 2 * w * (alpha * d - D) * alpha / np.where(d > 0, d, 1) / wd2, 0)
    g = c[:, None] * dv
    dy = np.zeros_like(y)
    np.add.at(dy, pi, g)
    np.add.at(dy, pj, -g)
    return L, dy


def loss_grad(params, off, nb, x, tgt, anchors, emulate_bf16=True):
    """Returns (loss, flat gradient) in the GPU's parameter layout."""
    r = bf16 if emulate_bf16 else (lambda v: np.asarray(v, np.float64))
    layers, Wh, bh = unpack(params)
    y, cache = forward(params, off, nb, x, emulate_bf16)
    pi, pj = pairs(off, nb, anchors)
    L, dy = loss_and_dy(y, tgt, pi, pj)
    grads = []
    Hlast = cache[-1][1]
    dWh = Hlast.T @ dy
    dbh = dy.sum(0)
    dH = dy @ Wh.T
    for li in range(len(layers) - 1, -1, -1):
        X, Hn, W = cache[li]
        dZ = np.where(Hn > 0, dH, 0.0)
        dW = X.T @ dZ
        db = dZ.sum(0)
        grads.append((dW, db))
        if li == 0:
            break
        dX = r(dZ) @ r(W).T
        F = X.shape[1] // 2
        dH = dX[:, :F] + mean_agg_t(off, nb, dX[:, F:])
    grads = grads[::-1]
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads] +
                          [dWh.ravel(), dbh])
    return L, flat
