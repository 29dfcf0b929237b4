"""Weighted CSR cases for the SpMV tests, in the shapes of the reference's own
tests (crates/graph-compute/tests/gpu_spmv.rs, gpu_spmv_hybrid.rs): the identity,
the 3x3 matrix whose product is [5, 6, 13], the f16 cases, random sparse
matrices of 0-4 entries per row from the tests' xorshift generator, and
power-law matrices whose hub rows touch every column."""
import numpy as np

MASK = (1 << 64) - 1


def _xorshift(seed: int):
    s = (seed + 0x9E3779B97F4A7C15) & MASK

    def nxt():
        nonlocal s
        s ^= (s << 13) & MASK
        s ^= s >> 7
        s ^= (s << 17) & MASK
        return s
    return nxt


def _unit(v: int) -> np.float32:
    # ((next % 2000) as f32) / 1000.0 - 1.0 in f32
    return np.float32(np.float32(v % 2000) / np.float32(1000.0)) - np.float32(1.0)


def identity(n: int):
    off = np.arange(n + 1, dtype=np.uint32)
    return off, np.arange(n, dtype=np.uint32), np.ones(n, np.float32)


def small_known():
    """A = [[2,0,1],[0,3,0],[1,0,4]], x = [1,2,3] -> y = [5,6,13]."""
    return (np.array([0, 2, 3, 5], np.uint32), np.array([0, 2, 1, 0, 2], np.uint32),
            np.array([2, 1, 3, 1, 4], np.float32), np.array([1, 2, 3], np.float32))


def f16_fractional():
    """gpu_spmv.rs f16_matches_f32_within_tolerance: values not exact in f16."""
    off = np.array([0, 2, 4, 6, 8], np.uint32)
    nb = np.array([1, 3, 0, 2, 1, 3, 0, 2], np.uint32)
    w = np.array([0.1, 0.7, 0.3333, 0.9, 0.123, 0.456, 0.789, 0.246], np.float32)
    x = np.array([0.31, 1.27, -0.55, 2.01], np.float32)
    return off, nb, w, x


def random_sparse(n: int, seed: int):
    """gpu_equals_cpu_spmv's matrix: 0..=4 entries per row, weights and x in [-1, 1)."""
    nxt = _xorshift(seed)
    off, nb, w = [0], [], []
    for _ in range(n):
        deg = nxt() % 5
        for _ in range(deg):
            nb.append(nxt() % n)
            w.append(_unit(nxt()))
        off.append(len(nb))
    x = [_unit(nxt()) for _ in range(n)]
    return (np.array(off, np.uint32), np.array(nb, np.uint32), np.array(w, np.float32),
            np.array(x, np.float32))


def power_law(n: int, hubs: int, seed: int):
    """gpu_spmv_hybrid.rs power_law: `hubs` rows adjacent to every node, the rest
    0..=3 random entries."""
    nxt = _xorshift(seed)
    off, nb, w = [0], [], []
    for v in range(n):
        if v < hubs:
            for c in range(n):
                nb.append(c)
                w.append(_unit(nxt()))
        else:
            for _ in range(nxt() % 4):
                nb.append(nxt() % n)
                w.append(_unit(nxt()))
        off.append(len(nb))
    x = [_unit(nxt()) for _ in range(n)]
    return (np.array(off, np.uint32), np.array(nb, np.uint32), np.array(w, np.float32),
            np.array(x, np.float32))


def ragged_rows(lengths, n: int, seed: int):
    """A square n x n matrix whose first rows have the given lengths (the rest
    empty), random columns, weights and x in [-1, 1)."""
    rng = np.random.default_rng(seed)
    lengths = list(lengths) + [0] * (n - len(lengths))
    off = np.zeros(n + 1, np.uint32)
    off[1:] = np.cumsum(lengths)
    nb = rng.integers(0, n, int(off[-1])).astype(np.uint32)
    w = rng.uniform(-1, 1, int(off[-1])).astype(np.float32)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    return off, nb, w, x


def order_sensitive_row():
    """One row of 65 entries whose serial and hybrid sums round differently.
    Products p_0 = 1, p_1 = 2^24, p_64 = 1, the rest 0 (x = 1):
      serial: (1 + 2^24) rounds to 2^24 (tie to even), + ... + 1 again -> 2^24;
      hybrid: lane 0 holds 1 + 1 = 2, lane 1 holds 2^24, the tree's last step
              adds them: 2^24 + 2, exact."""
    w = np.zeros(65, np.float32)
    w[0], w[1], w[64] = 1.0, 2.0 ** 24, 1.0
    off = np.array([0, 65], np.uint32)
    nb = np.zeros(65, np.uint32)
    return off, nb, w, np.ones(1, np.float32)
