"""Adversarial inputs of the reference's engine fuzz suite
(crates/graph-compute/tests/fuzz.rs:65-237): graph shapes (path, hub star,
scale-free, raw pairs with self-loops and duplicates, islands with isolated
nodes, no edges), seed positions (ball, coincident, 1e-4 cluster, collinear,
1e5 magnitudes, zeros) and the fa2 settings ranges, drawn from SplitMix64
(stability.rs:11-26) with fixed seeds so every run sees the same cases."""
import numpy as np

from oracle import oracle as O

_MASK = (1 << 64) - 1
SHAPES = ["path", "star", "scale_free", "raw_pairs", "islands", "edge_free"]
SEEDS = ["ball", "coincident", "epsilon_cluster", "collinear", "huge", "zeros"]


class SplitMix64:
    def __init__(self, state):
        self.s = state & _MASK

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & _MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def next_f32(self):
        return np.float32(self.next_u64() >> 40) / np.float32(1 << 24)

    def uniform(self, lo, hi):
        return np.float32(lo + (hi - lo) * float(self.next_f32()))


def _csr(n, pairs):
    return O.csr_from_pairs(n, np.array(pairs, np.uint32).reshape(-1))


def build_graph(shape, n, seed):
    """fuzz.rs build_graph (:133-145)."""
    if shape == "path":
        return _csr(n, [[v - 1, v] for v in range(1, n)])
    if shape == "star":
        return _csr(n, [[0, v] for v in range(1, n)])
    if shape == "scale_free":
        return O.ba_graph(max(n, 2), 3, seed)
    if shape == "raw_pairs":  # fuzz.rs:68-81: self-loops and duplicates on purpose
        rng = SplitMix64(seed)
        return _csr(n, [[v, rng.next_u64() % n] for v in range(n) for _ in range(4)])
    if shape == "islands":  # fuzz.rs:91-104: two trees, then isolated nodes
        half = max(n // 3, 1)
        rng = SplitMix64(seed)
        pairs = [[v, rng.next_u64() % v] for v in range(1, half)]
        pairs += [[v, half + rng.next_u64() % (v - half)] for v in range(half + 1, 2 * half)]
        return _csr(n, pairs)
    return _csr(n, [])


def build_seed(kind, n, seed):
    """fuzz.rs build_seed (:188-221)."""
    rng = SplitMix64(seed)
    f = rng.next_f32
    if kind == "ball":
        return O.ball_seed(n, seed)
    if kind == "coincident":
        x, y, z = (f() * np.float32(100.0) - np.float32(50.0) for _ in range(3))
        return np.tile(np.array([x, y, z], np.float32), n)
    if kind == "epsilon_cluster":
        return np.array([f() * np.float32(1e-4) for _ in range(3 * n)], np.float32)
    if kind == "collinear":
        p = np.zeros(3 * n, np.float32)
        p[0::3] = np.arange(n, dtype=np.float32) * np.float32(0.5)
        return p
    if kind == "huge":
        return np.array([(f() * np.float32(2.0) - np.float32(1.0)) * np.float32(1e5) for _ in range(3 * n)],
                        np.float32)
    return np.zeros(3 * n, np.float32)


def fa2_params(rng):
    """fuzz.rs fa2_params (:238-262): max_displacement stays > 0."""
    return dict(gravity=rng.uniform(0.0, 20.0), strong_gravity=bool(rng.next_u64() & 1),
                scaling_ratio=rng.uniform(0.01, 20.0), edge_weight_influence=rng.uniform(0.0, 2.0),
                jitter_tolerance=rng.uniform(0.01, 10.0), lin_log_mode=bool(rng.next_u64() & 1),
                time_step=rng.uniform(0.01, 2.0), max_displacement=rng.uniform(1.0, 50.0))


def gpu_n(rng):
    """fuzz.rs gpu_n (:170-176): 1, 8..512 or 513..4096 (weights 1 : 3 : 2)."""
    r = rng.next_u64() % 6
    if r == 0:
        return 1
    if r < 4:
        return 8 + rng.next_u64() % 505
    return 513 + rng.next_u64() % 3584


def cases(seed=0xF022, count=36):
    """Every shape with every seed kind once (36 cases); n, graph seed, steps
    (5..25) and the fa2 settings drawn per case."""
    rng = SplitMix64(seed)
    out = []
    for c in range(count):
        shape, kind = SHAPES[c % 6], SEEDS[(c // 6) % 6]
        n = int(gpu_n(rng))
        gseed = rng.next_u64()
        steps = 5 + int(rng.next_u64() % 21)
        out.append(dict(case=c, shape=shape, seed_kind=kind, n=n, graph_seed=gseed, steps=steps,
                        fa2=fa2_params(rng)))
    return out


def ceiling(seed_xyz, steps, step_cap):
    """fuzz.rs check_engine (:307-312): seed extent + (steps + 1) * cap * 2."""
    seed_max = float(np.abs(seed_xyz).max()) if len(seed_xyz) else 0.0
    return seed_max + (steps + 1.0) * step_cap * 2.0
