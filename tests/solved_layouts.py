"""Graphs and seeds of the reference's solved-layout canaries
(crates/graph-compute/tests/solved_layouts.rs), shared by the CPU oracle pins
and the GPU quality test."""
import numpy as np

from oracle import oracle as O

_MASK = (1 << 64) - 1


def seed(n, spread):
    """solved_layouts.rs `seed`: x and y from the SplitMix64 finaliser of
    2i+1 and 2i+7 (plus the golden increment), (z / u64::MAX as f32 * 2 - 1) *
    spread; z = 0."""
    def h(k):
        z = (k + 0x9E3779B97F4A7C15) & _MASK
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        z ^= z >> 31
        x = np.float32(z / float(_MASK))
        return (x * np.float32(2.0) - np.float32(1.0)) * np.float32(spread)
    p = np.zeros(3 * n, np.float32)
    for i in range(n):
        p[3 * i] = h(2 * i + 1)
        p[3 * i + 1] = h(2 * i + 7)
    return p


def _csr(n, pairs):
    return O.csr_from_pairs(n, np.array(pairs, np.uint32).reshape(-1))


def path(n):
    return _csr(n, [[i, i + 1] for i in range(n - 1)])


def cycle(n):
    return _csr(n, [[i, (i + 1) % n] for i in range(n)])


def grid(w, h):
    pairs = []
    for r in range(h):
        for c in range(w):
            v = r * w + c
            if c + 1 < w:
                pairs.append([v, v + 1])
            if r + 1 < h:
                pairs.append([v, v + w])
    return _csr(w * h, pairs)


def complete(n):
    return _csr(n, [[i, j] for i in range(n) for j in range(i + 1, n)])


CASES = {"path24": lambda: path(24), "cycle16": lambda: cycle(16), "grid4x4": lambda: grid(4, 4),
         "complete6": lambda: complete(6)}


def normalized_stress(pos, off, nb):
    """solved_layouts.rs scale_normalized_stress: all reachable pairs i < j."""
    ti, tj, td = O.all_pairs_terms(off, nb)
    return O.stress_normalized(pos, ti, tj, td)
