"""The learned warm start on the GPU: the committed network, run by the
library's own kernels, against the plain FA2+BH engine from the same random
start on a held-out Barabasi-Albert graph (not among the training graphs:
other sizes and seeds).  The gain is measured the way the headline metric is:
time to reach normalized stress <= 1.01 x the plain run's converged stress,
with the plain run as the fallback; the ablation starts the plain engine from
the same start rescaled to the size law, without the network."""
import numpy as np
import pytest

import jumpcannon as jc
from jumpcannon import warmstart as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net():
    return W.load()


def test_weights_load_and_match_the_network(net):
    m, coef = net
    assert m.params().shape == (jc.Mpnn(1).P,)
    assert np.all(np.isfinite(m.params())) and coef.shape == (3,)


def test_warm_start_beats_plain_and_the_scale_only_ablation(net):
    m, coef = net
    off, nb = jc.ba_graph(40_000, 4, 0xBADC0DE)
    x0 = jc.uniform_cube(40_000, 10.0, 0xBADC0DE + 1)
    r = W.evaluate(m, coef, off, nb, x0, horizon=1200, every=50, theta=0.8)
    print({k: v for k, v in r.items() if not k.endswith("curve")})
    assert not r["fallback_to_plain"]
    assert r["learned_init_stress"] < r["plain_init_stress"]
    assert r["speedup"] >= 3.0
    sr = r["scaled_random_ms_to_target"]
    assert sr is None or r["learned_ms_to_target"] < sr


def test_prediction_is_deterministic_and_finite(net):
    m, coef = net
    off, nb = jc.ba_graph(5_000, 3, 77)
    x0 = jc.uniform_cube(5_000, 10.0, 78)
    a, _, _ = W.predict(m, coef, off, nb, x0)
    b, _, _ = W.predict(m, coef, off, nb, x0)
    assert np.isfinite(a).all()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_device_rescale_matches_the_host_f64_rescale(net):
    """forward_scaled centres and scales on the device in f64; the host form of
    the same arithmetic (numpy f64 mean, RMS radius, scale, cast to f32) agrees
    to within one f32 rounding of each coordinate, the mean is ~0 and the RMS
    radius is the target."""
    m, coef = net
    off, nb = jc.ba_graph(20_000, 4, 91)
    x0 = jc.uniform_cube(20_000, 10.0, 92)
    m.set_graph(off, nb, feat_seed=7, walks=16, noise=np.asarray(x0) / W.INIT_HALF)
    y, _ = m.forward()
    target = 37.5
    got, _ = m.forward_scaled(target)
    p = y.reshape(-1, 3).astype(np.float64)
    p -= p.mean(0)
    want = (p * (target / np.sqrt((p ** 2).sum(1).mean()))).astype(np.float32).reshape(-1)
    tol = 2.0 * np.spacing(np.abs(want).astype(np.float32)) + 1e-6
    assert np.all(np.abs(got - want) <= tol)
    g = got.reshape(-1, 3).astype(np.float64)
    assert np.abs(g.mean(0)).max() < 1e-4 * target
    assert abs(np.sqrt((g ** 2).sum(1).mean()) - target) < 1e-4 * target


def test_warm_start_on_the_held_out_1m_graph(net):
    """The held-out 1M-node BA workload graph (seed 0xC0FFEE; training used other
    seeds and at most 20k nodes), the bench's start and theta: the learned start
    reaches stress <= 1.01 S* at least 3x sooner than plain FA2 from random (the
    stress criterion), and after the same refinement its layout is of plain
    quality by both judges (stress <= 1.01 S*, edge-length CV within 1 % of
    plain's converged CV).  The time to bring both judges within 1 % is printed:
    there the learned start is not faster (its edge lengths are FA2's only after
    refinement)."""
    from jumpcannon import workloads

    m, coef = net
    off, nb, xyz = workloads.ba_1m()
    r = W.evaluate(m, coef, off, nb, xyz, horizon=2000, every=50, theta=0.8, ablation=False, multilevel=False)
    keys = ("plain_stress_final", "plain_cv_final", "learned_init_stress", "learned_init_cv",
            "plain_ms_to_target", "learned_ms_to_target", "speedup", "learned_max_stress_after_target",
            "plain_ms_to_both", "learned_ms_to_both", "speedup_both", "learned_stays_within_both",
            "learned_stress_final", "learned_cv_final")
    print({k: r[k] for k in keys})
    assert not r["fallback_to_plain"] and r["speedup"] >= 3.0
    assert r["learned_stress_final"] <= r["target_stress"]
    assert abs(r["learned_cv_final"] - r["plain_cv_final"]) <= 0.01 * r["plain_cv_final"]


def test_training_stream_resumes_exactly():
    """Stopping the training stream after graph 3 and resuming from the saved
    parameters and Adam state gives the same parameters, bit for bit, as one
    uninterrupted pass over graphs 0..5 (small graphs, 600-2,000 nodes)."""
    kw = dict(total=6, seed=777, n_min=600, n_max=2000, pivots=8, sgd_epochs=6)
    a = jc.Mpnn(seed=5)
    la = W.train_stream(a, 0, 6, **kw)
    b = jc.Mpnn(seed=5)
    lb = W.train_stream(b, 0, 3, **kw)
    m1, m2, step = b.adam_state()
    assert step == 6
    c = jc.Mpnn(seed=11)
    c.set_params(b.params())
    c.set_adam_state(m1, m2, step)
    lb += W.train_stream(c, 3, 6, **kw)
    assert np.array_equal(la, lb)
    assert np.array_equal(a.params(), c.params())
