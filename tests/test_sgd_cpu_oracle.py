"""Pins of the oracle's restatement of the reference CPU s_gd2 engine
(SgdStressEngine, crates/graph-compute/src/engines/sgd_stress.rs:251-366): the
reference's own CPU canaries (solved_layouts.rs) and a closed form whose
sampling count comes from the published SplitMix64 stream."""
import math

import numpy as np

from oracle import oracle as O
import solved_layouts as S


def test_path_zero_stress_fixed_point_cpu():
    """solved_layouts.rs:254-282: the unit line has ~0 stress, and a perturbed
    line relaxes back (NS < 0.05 after 150 steps)."""
    n = 20
    off, nb = S.path(n)
    line = np.zeros(3 * n, np.float32)
    line[0::3] = np.arange(n, dtype=np.float32)
    assert S.normalized_stress(line, off, nb) < 1e-4
    init = line + S.seed(n, 0.4)
    out = O.sgd_cpu_run(off, nb, init, 150)
    ns = S.normalized_stress(out, off, nb)
    print(f"path20 NS {S.normalized_stress(init, off, nb):.4f} -> {ns:.5f}")
    assert ns < 0.05


def test_cycle_stays_roughly_regular_cpu():
    """solved_layouts.rs:359-392: a 16-cycle seeded as a polygon of radius 5 plus
    noise stays regular (radius CV < 0.25 after 150 steps)."""
    n = 16
    off, nb = S.cycle(n)
    init = np.zeros(3 * n, np.float32)
    noise = S.seed(n, 0.15)
    for i in range(n):
        t = np.float32(i) / np.float32(n) * np.float32(2 * math.pi)
        init[3 * i] = np.float32(math.cos(t)) * np.float32(5.0) + noise[3 * i]
        init[3 * i + 1] = np.float32(math.sin(t)) * np.float32(5.0) + noise[3 * i + 1]
    p = O.sgd_cpu_run(off, nb, init, 150).reshape(-1, 3)
    r = np.linalg.norm(p[:, :2] - p[:, :2].mean(0), axis=1)
    assert r.std() / r.mean() < 0.25


def test_pair_update_closed_form_and_sampling_count():
    """Two nodes, one edge, one pivot, eta held at eta_min (no annealing): the
    pivot is draw 0 mod 2 and each of the n = 2 samples per sweep is the pair
    (draw mod 2, draw mod 1); every sample whose node is not the pivot moves both
    endpoints half of mu (|D| - 1) toward each other.  So after u updates
    |D_u| - 1 = (|D_0| - 1)(1 - mu)^u, and the midpoint never moves."""
    off, nb = S.path(2)
    init = np.array([0, 0, 0, 3.0, 4.0, 0], np.float32)  # |D_0| = 5
    seed, steps, eta = 12345, 7, 0.3
    draws = [int(x) for x in O.splitmix_u64(seed, 1 + 2 * 2 * steps)]
    pivot = draws[0] % 2
    u = sum(1 for c in range(2 * steps) if draws[1 + 2 * c] % 2 != pivot)
    assert 0 < u < 2 * steps
    out = O.sgd_cpu_run(off, nb, init, steps, n_pivots=1, eta_max=1.0, eta_min=eta, n_anneal=0,
                        seed=seed).reshape(2, 3).astype(np.float64)
    d = np.linalg.norm(out[1] - out[0])
    want = 1.0 + 4.0 * (1.0 - eta) ** u
    assert abs(d - want) < 1e-5 * want, (d, want, u)
    assert np.allclose((out[0] + out[1]) / 2, [1.5, 2.0, 0.0], atol=1e-5)
    assert np.allclose((out[1] - out[0]) / d, [0.6, 0.8, 0.0], atol=1e-5)
