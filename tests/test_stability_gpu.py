"""Whole-run stability of the device engines: the reference's vault-scale
regression (crates/graph-compute/tests/layout_stability.rs:91-150) and its
engine fuzz suite (tests/fuzz.rs), here with fixed cases, each also compared
step by step with the oracle's restatement of the reference path."""
import numpy as np
import pytest

import jumpcannon as jc
from oracle import oracle as O
import fuzz_cases as F
from stability_rules import assert_stable

pytestmark = pytest.mark.gpu


def _run_collect(engine, off, nb, steps, seed=0x5EED):
    """layout_stability.rs run_collect (:45-68): ball seed, StepStats per step,
    no non-finite coordinate at any step."""
    n = len(off) - 1
    prev = jc.ball_seed(n, seed)
    engine.init(off, nb, prev)
    stats = []
    for step in range(1, steps + 1):
        out = engine.step(1)
        st = jc.step_stats(prev, out).as_dict()
        assert st["nonfinite"] == 0, f"non-finite positions at step {step}"
        stats.append(st)
        prev = out
    return stats


def test_fa2_bh_stable_at_vault_scale_synthetic():
    """layout_stability.rs:131-146: synthetic_scale_free(9724, 5, 0xC0FFEE), ball
    seed, default settings, 1,000 steps."""
    off, nb = jc.ba_graph(9724, 5, 0xC0FFEE)
    n = len(off) - 1
    stats = _run_collect(jc.Fa2BhEngine(), off, nb, 1000)
    early, mid, late = assert_stable("fa2-bh-hip/synthetic-10k", stats, np.sqrt(n) * 5.0, 10.0)
    print(f"vault-scale: p50 disp early {early:.4f} mid {mid:.4f} late {late:.4f}; "
          f"final p50 radius {stats[-1]['p50_radius']:.1f}, max |pos| {stats[-1]['max_abs']:.1f}")


def test_fa2_brute_stable_on_synthetic_scale_free():
    """layout_stability.rs:150-165: fa2-brute on synthetic_scale_free(2000, 5,
    0xC0FFEE), ball seed, 700 steps."""
    off, nb = jc.ba_graph(2000, 5, 0xC0FFEE)
    n = len(off) - 1
    stats = _run_collect(jc.Fa2BruteEngine(), off, nb, 700)
    assert_stable("fa2-brute-hip/synthetic-2k", stats, np.sqrt(n) * 5.0, 10.0)


def test_theta0_stable_on_synthetic_scale_free():
    """The same graph and bar through the Barnes-Hut walk at theta = 0 (every
    leaf visited, the brute-force repulsion in tree order)."""
    off, nb = jc.ba_graph(2000, 5, 0xC0FFEE)
    n = len(off) - 1
    stats = _run_collect(jc.Fa2BhEngine(theta=0.0), off, nb, 700)
    assert_stable("fa2-bh-hip theta=0/synthetic-2k", stats, np.sqrt(n) * 5.0, 10.0)


CASES = F.cases()


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"case{c['case']}-{c['shape']}-{c['seed_kind']}-n{c['n']}")
def test_fuzz_fa2_bh_hip(case):
    """fuzz.rs:400-447 (gpu_fuzz fa2-bh): init accepts the input; every step
    returns 3n finite coordinates under the kinematic ceiling.  Also every
    step equals the oracle bit for bit (lin-log cases excepted: their logf
    is only pinned to a tolerance)."""
    off, nb = F.build_graph(case["shape"], case["n"], case["graph_seed"])
    n = len(off) - 1
    seed = F.build_seed(case["seed_kind"], n, case["graph_seed"] ^ 0xA5)
    p = case["fa2"]
    e = jc.Fa2BhEngine()
    e.set_params({k: (bool(v) if isinstance(v, bool) else float(v)) for k, v in p.items()})
    e.init(off, nb, seed)
    cap = float(p["max_displacement"]) * max(float(p["time_step"]), 1.0)
    ceil = F.ceiling(seed, case["steps"], cap)
    s = O.Fa2Settings.make(**{k: (int(v) if isinstance(v, bool) else float(v)) for k, v in p.items()})
    x, old, se = seed, None, None
    for step in range(1, case["steps"] + 1):
        out = e.step(1)
        assert len(out) == 3 * n and np.isfinite(out).all(), step
        assert np.abs(out).max() <= ceil, (step, float(np.abs(out).max()), ceil)
        if not p["lin_log_mode"]:
            x, old, se, _ = O.fa2_run(off, nb, x, 1, s, tree_kind=1, exact=True, old_force=old, se=se)
            assert np.array_equal(out.view(np.uint32), x.view(np.uint32)), step


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"case{c['case']}-{c['shape']}-{c['seed_kind']}-n{c['n']}")
def test_fuzz_fa2_brute_hip(case):
    """fuzz.rs:400-445 (gpu_fuzz fa2-brute): the same cases and invariants
    through fa2-brute-hip, and every step equal to the oracle's brute-force
    iteration bit for bit (lin-log cases excepted)."""
    off, nb = F.build_graph(case["shape"], case["n"], case["graph_seed"])
    n = len(off) - 1
    seed = F.build_seed(case["seed_kind"], n, case["graph_seed"] ^ 0xA5)
    p = case["fa2"]
    e = jc.Fa2BruteEngine()
    e.set_params({k: (bool(v) if isinstance(v, bool) else float(v)) for k, v in p.items()})
    e.init(off, nb, seed)
    cap = float(p["max_displacement"]) * max(float(p["time_step"]), 1.0)
    ceil = F.ceiling(seed, case["steps"], cap)
    s = O.Fa2Settings.make(**{k: (int(v) if isinstance(v, bool) else float(v)) for k, v in p.items()})
    x, old, se = seed, None, None
    for step in range(1, case["steps"] + 1):
        out = e.step(1)
        assert len(out) == 3 * n and np.isfinite(out).all(), step
        assert np.abs(out).max() <= ceil, (step, float(np.abs(out).max()), ceil)
        if not p["lin_log_mode"]:
            x, old, se, _ = O.fa2_run(off, nb, x, 1, s, tree_kind=2, exact=True, old_force=old, se=se)
            assert np.array_equal(out.view(np.uint32), x.view(np.uint32)), step


@pytest.mark.parametrize("case", CASES[:18], ids=lambda c: f"case{c['case']}-{c['shape']}-{c['seed_kind']}-n{c['n']}")
def test_fuzz_sgd_stress_hip(case):
    """fuzz.rs:368-384 (sgd-stress) through sgd-stress-hip: finite, 3n, under
    the seed extent + 100 ceiling; pivots and every step equal the oracle."""
    off, nb = F.build_graph(case["shape"], case["n"], case["graph_seed"])
    n = len(off) - 1
    seed = F.build_seed(case["seed_kind"], n, case["graph_seed"] ^ 0xA5)
    e = jc.SgdStressEngine()
    e.init(off, nb, seed)
    piv, dist, eta_max = e.pivots()
    opiv, odist = O.select_pivots(off, nb, 50)
    assert np.array_equal(piv, opiv) and np.array_equal(dist, odist)
    ceil = F.ceiling(seed, case["steps"], 100.0 + float(np.abs(seed).max()))
    pos = seed
    for step in range(1, case["steps"] + 1):
        out = e.step(1)
        assert len(out) == 3 * n and np.isfinite(out).all() and np.abs(out).max() <= ceil, step
        pos = O.sgd_sweep(piv, dist, pos, O.anneal_eta(step - 1, eta_max, 0.01, 30), threads=8)
        assert np.array_equal(out.view(np.uint32), pos.view(np.uint32)), step
