"""CPU checks of the stability fixtures and telemetry (stability.rs:30-50,
128-186) in the product, and of the fuzz suite's invariants on the oracle's
restatement of the reference path (fuzz.rs:290-330)."""
import numpy as np
import pytest

import jumpcannon as jc
from oracle import oracle as O
import fuzz_cases as F


@pytest.mark.parametrize("n,seed", [(1, 0), (7, 0x5EED), (9724, 0x5EED)])
def test_ball_seed_equals_oracle_and_stays_in_the_ball(n, seed):
    p = jc.ball_seed(n, seed)
    assert np.array_equal(p.view(np.uint32), O.ball_seed(n, seed).view(np.uint32))
    r = np.linalg.norm(p.reshape(-1, 3).astype(np.float64), axis=1)
    assert r.max() <= np.sqrt(n) * 5.0 * (1 + 1e-6)


def _stats_numpy(prev, cur):
    """StepStats::measure restated with numpy's sort and IEEE f32 elementwise
    operations in the reference's order."""
    c = cur.reshape(-1, 3)
    p = prev.reshape(-1, 3)
    ok = np.isfinite(c).all(axis=1)
    cf, pf = c[ok], p[ok]
    radii = np.sort(np.sqrt(cf[:, 0] * cf[:, 0] + cf[:, 1] * cf[:, 1] + cf[:, 2] * cf[:, 2]))
    d = cf - pf
    dd = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]).astype(np.float64))
    disps = np.sort(dd.astype(np.float32))
    pick = lambda v, q: float(v[int((len(v) - 1) * q)]) if len(v) else 0.0
    n = len(c)
    return {"max_abs": float(np.abs(cf).max()) if len(cf) else 0.0, "p50_radius": pick(radii, 0.5),
            "p99_radius": pick(radii, 0.99), "mean_disp": float(np.float32(dd.sum() / n)) if n else 0.0,
            "p50_disp": pick(disps, 0.5), "nonfinite": int((~ok).sum())}


def test_step_stats_against_numpy():
    rng = np.random.default_rng(4)
    prev = (rng.standard_normal(3 * 1001) * 50).astype(np.float32)
    cur = prev + (rng.standard_normal(3 * 1001) * 2).astype(np.float32)
    cur[30], cur[304], cur[3000] = np.nan, np.inf, -np.inf
    got = jc.step_stats(prev, cur).as_dict()
    want = _stats_numpy(prev, cur)
    assert got["nonfinite"] == 3
    for k, v in want.items():
        if k == "mean_disp":  # the f64 sum runs in another order
            assert abs(got[k] - v) <= 1e-6 * v, k
        else:
            assert got[k] == v, k
    z = jc.step_stats(np.zeros(0, np.float32), np.zeros(0, np.float32)).as_dict()
    assert z == {"max_abs": 0.0, "p50_radius": 0.0, "p99_radius": 0.0, "mean_disp": 0.0, "p50_disp": 0.0,
                 "nonfinite": 0}
    with pytest.raises(jc.JcError):
        jc.step_stats(np.zeros(3, np.float32), np.zeros(6, np.float32))


def _oracle_steps(off, nb, xyz, steps, settings):
    """fa2_run one iteration at a time: the positions after every step."""
    old, se = None, None
    out = []
    for _ in range(steps):
        xyz, old, se, _ = O.fa2_run(off, nb, xyz, 1, settings, tree_kind=1, exact=True, old_force=old, se=se)
        out.append(xyz)
    return out


@pytest.mark.parametrize("case", [c for c in F.cases() if c["n"] <= 600][:10], ids=lambda c: f"case{c['case']}")
def test_fuzz_invariants_on_the_oracle(case):
    """fuzz.rs:290-330 on the restated reference path (small cases; the device
    runs all of them against this oracle)."""
    off, nb = F.build_graph(case["shape"], case["n"], case["graph_seed"])
    n = len(off) - 1
    seed = F.build_seed(case["seed_kind"], n, case["graph_seed"] ^ 0xA5)
    p = case["fa2"]
    s = O.Fa2Settings.make(**{k: (int(v) if isinstance(v, bool) else float(v)) for k, v in p.items()})
    cap = float(p["max_displacement"]) * max(float(p["time_step"]), 1.0)
    ceil = F.ceiling(seed, case["steps"], cap)
    for k, x in enumerate(_oracle_steps(off, nb, seed, case["steps"], s)):
        assert len(x) == 3 * n and np.isfinite(x).all(), k
        assert np.abs(x).max() <= ceil, k
