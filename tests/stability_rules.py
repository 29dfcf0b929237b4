"""The reference's whole-run stability rules (crates/graph-compute/tests/
layout_stability.rs:74-129): no breach of the kinematic ceiling, a bulk median
radius that neither collapses nor runs away, and a median displacement that
decays across the run instead of oscillating."""
import numpy as np


def window_disp(stats, lo, hi):
    v = sorted(s["p50_disp"] for s in stats[lo:hi])
    return v[len(v) // 2]


def assert_stable(name, stats, seed_radius, max_displacement):
    steps = len(stats)
    last = stats[-1]
    ceiling = seed_radius + steps * max_displacement
    assert last["max_abs"] <= 1.01 * ceiling, f"{name}: max |pos| {last['max_abs']} breaches {ceiling}"
    assert 1.0 < last["p50_radius"] < 0.25 * ceiling, f"{name}: bulk median radius {last['p50_radius']}"
    early = window_disp(stats, steps // 10, steps // 3)
    mid = window_disp(stats, steps // 3, 2 * steps // 3)
    late = window_disp(stats, 8 * steps // 10, steps)
    assert late <= mid * 1.15 and mid <= early * 1.15, f"{name}: not decaying {early} {mid} {late}"
    assert late < 0.5 * early or late < 0.05, f"{name}: barely decayed {early} -> {late}"
    return early, mid, late
