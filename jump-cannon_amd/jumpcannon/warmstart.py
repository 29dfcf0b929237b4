"""Learned warm start for FA2+BH: train the message-passing network on layouts
the plain engine produces, predict initial coordinates, refine with the plain
engine, and judge both runs by normalized pivot stress.

The plain method stays the drop-in and the fallback: a prediction whose
refinement does not reach the plain run's target stress within the plain run's
own iteration budget is replaced by the plain run.

Training sample: (graph, random initial layout x0, target layout).  The target
keeps what the network can know and the stress engine decides: each node's
distance from the centre of the layout the project's SGD sparse-stress engine
reaches from x0 (max-min pivots, BFS hop distances, annealed sweeps), along the
direction of the node's own noise averaged over its 3-hop neighbourhood (the
network's receptive field; the stress layout's global orientation of a region
is not predictable from 3 hops).  On Barabasi-Albert graphs this radial profile
is also where the normalized stress lives: with random directions it already
scores lower than the stress layout it came from.  The loss is invariant to
rotation, translation and scale; the scale comes from a size law fitted on the
converged plain FA2 layouts of the same graphs, so the prediction starts at the
size FA2 settles at.  The network sees x0/10 as its noise input.
"""
from __future__ import annotations

import math
import os
import time

import numpy as np

from . import (Fa2BhBatch, Fa2BhEngine, Mpnn, MultilevelEngine, SgdStressEngine, StressJudge, _splitmix_pair,
               ba_graph, sbm_graph, uniform_cube)

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warmstart_weights.npz")
INIT_HALF = 10.0     # initial layouts are uniform in [-10,10]^3 (the 1M-node workload's seed)
TRAIN_PIVOT_SEED = 0x5EEDC0DE12345678   # sgd-stress default seed: the training targets
JUDGE_PIVOT_SEED = 0x4A554447455F3031   # a different pivot set judges the layouts


def training_graph(g: int, seed: int, n_min: int = 1000, n_max: int = 20000):
    """Graph g of the training stream: SBM (even g) or BA (odd g) of log-uniform
    size in [n_min, n_max], with its random initial layout x0."""
    u, h = _splitmix_pair(seed + g)
    n = int(n_min * (n_max / n_min) ** u)
    if g % 2 == 0:
        B = 2 + h % 7
        off, nb = sbm_graph(n, B, min(1.0, 8.0 / (n / B)), 0.5 / n, seed + g)
    else:
        off, nb = ba_graph(n, 2 + h % 4, seed + g)
    x0 = uniform_cube(len(off) - 1, INIT_HALF, seed + 100_000 + g)
    return off, nb, x0


def training_graphs(count: int, seed: int, n_min: int = 1000, n_max: int = 20000, start: int = 0):
    return [training_graph(g, seed, n_min, n_max) for g in range(start, start + count)]


def fa2_targets(graphs, iters: int = 1500, theta: float = 1.0):
    """Plain FA2 layouts of every graph from its x0, all graphs in one batch."""
    b = Fa2BhBatch(theta=theta)
    b.init([(o, nb) for o, nb, _ in graphs], [x0 for _, _, x0 in graphs])
    return b.step(iters)


def smooth_directions(off, nb, x0, hops: int = 3):
    """Unit vectors of the centred x0 averaged `hops` times over each node and
    its neighbours."""
    off = np.asarray(off, np.int64)
    nb = np.asarray(nb, np.int64)
    deg = np.diff(off)
    n = len(deg)
    src = np.repeat(np.arange(n), deg)
    u = np.asarray(x0, np.float64).reshape(-1, 3)
    u = u - u.mean(0)
    for _ in range(hops):
        acc = np.empty_like(u)
        for c in range(3):
            acc[:, c] = u[:, c] + np.bincount(src, weights=u[nb, c], minlength=n)
        u = acc / (deg + 1)[:, None]
    return u / np.maximum(np.linalg.norm(u, axis=1, keepdims=True), 1e-30)


def stress_target(off, nb, x0, k: int = 50, epochs: int = 30, seed: int = TRAIN_PIVOT_SEED):
    """The SGD sparse-stress layout of the graph from x0 (k max-min pivots, eta
    from (max hop)^2 down to 0.01 over `epochs` sweeps); the target is its radial
    profile along the smoothed noise directions."""
    e = SgdStressEngine(n_pivots=k, eta_max=0.0, eta_min=0.01, n_anneal_steps=epochs, seed=seed)
    e.init(off, nb, x0)
    sg = e.step(epochs).reshape(-1, 3).astype(np.float64)
    e.close()
    r = np.linalg.norm(sg - sg.mean(0), axis=1)
    return (smooth_directions(off, nb, x0) * r[:, None]).astype(np.float32).reshape(-1)


def stress_targets(graphs, k: int = 50, epochs: int = 30, seed: int = TRAIN_PIVOT_SEED):
    return [stress_target(off, nb, x0, k, epochs, seed) for off, nb, x0 in graphs]


def rms_radius(xyz):
    p = np.asarray(xyz, np.float64).reshape(-1, 3)
    return float(np.sqrt(((p - p.mean(0)) ** 2).sum(1).mean()))


def fit_size_law(graphs, targets):
    """log R_rms = a + b log n + c log(mean degree), least squares."""
    X, y = [], []
    for (off, nb, _), t in zip(graphs, targets):
        n = len(off) - 1
        X.append([1.0, math.log(n), math.log(max(len(nb) / n, 1e-9))])
        y.append(math.log(rms_radius(t)))
    coef, *_ = np.linalg.lstsq(np.array(X), np.array(y), rcond=None)
    return coef


def size_of(coef, n, nnz):
    return float(math.exp(coef[0] + coef[1] * math.log(n) + coef[2] * math.log(max(nnz / n, 1e-9))))


def train(net: Mpnn, graphs, targets, epochs: int = 4, steps_per_graph: int = 2,
          lr: float = 2e-3, n_anchor: int = 32, walks: int = 16, log=None):
    losses = []
    for ep in range(epochs):
        for g, ((off, nb, x0), t) in enumerate(zip(graphs, targets)):
            net.set_graph(off, nb, feat_seed=7 + g, walks=walks, noise=x0 / INIT_HALF)
            net.set_target(t, n_anchor=n_anchor, seed=ep * 1000 + g)
            losses.append(net.train(steps_per_graph, lr))
        if log:
            log(f"epoch {ep}: mean loss {np.mean(losses[-len(graphs):]):.4f}")
    return losses


def save(net: Mpnn, coef, path=WEIGHTS, **meta):
    np.savez(path, params=net.params(), size_law=np.asarray(coef, np.float64),
             **{k: np.asarray(v) for k, v in meta.items()})


def cosine_lr(i: int, total: int, lr: float, lr_min: float) -> float:
    """Learning rate of step i of `total`: cosine from lr down to lr_min."""
    return lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * min(i, total) / max(total, 1)))


def train_stream(net: Mpnn, start: int, stop: int, total: int, seed: int, n_min: int, n_max: int,
                 steps_per_graph: int = 2, lr: float = 2e-3, lr_min: float = 2e-4, n_anchor: int = 32,
                 walks: int = 16, pivots: int = 50, sgd_epochs: int = 30, log=None, every: int = 500):
    """One pass over graphs [start, stop) of a stream of `total`: each graph is
    generated, given its stress target, trained on for steps_per_graph Adam
    steps and dropped, so the stream's size does not bound memory.  Returns the
    per-graph losses."""
    losses = []
    for g in range(start, stop):
        off, nb, x0 = training_graph(g, seed, n_min, n_max)
        t = stress_target(off, nb, x0, pivots, sgd_epochs)
        net.set_graph(off, nb, feat_seed=7 + g, walks=walks, noise=x0 / INIT_HALF)
        net.set_target(t, n_anchor=n_anchor, seed=g)
        losses.append(net.train(steps_per_graph, cosine_lr(g, total, lr, lr_min)))
        if log and (g + 1 - start) % every == 0:
            log(f"graph {g + 1}/{total}: mean loss of the last {every} {np.mean(losses[-every:]):.4f}")
    return losses


def load(path=WEIGHTS, seed: int = 1):
    with np.load(path, allow_pickle=False) as z:
        net = Mpnn(seed)
        net.set_params(z["params"])
        return net, z["size_law"].copy()


def predict(net: Mpnn, coef, off, nb, x0, walks: int = 16, feat_seed: int = 7):
    """Predicted initial layout (3n) and the device time of features + forward (ms)."""
    t0 = time.perf_counter()
    net.set_graph(off, nb, feat_seed=feat_seed, walks=walks, noise=np.asarray(x0) / INIT_HALF)
    # centred and scaled to the size law's RMS radius on the device (f64 sums)
    p, fwd_ms = net.forward_scaled(size_of(coef, len(off) - 1, len(nb)))
    wall_ms = (time.perf_counter() - t0) * 1e3
    return p, fwd_ms, wall_ms


def stress_judge(off, nb, k: int = 50, seed: int = JUDGE_PIVOT_SEED):
    """Pivot terms of k max-min pivots (sgd_stress.rs:143-175) on the GPU stress plan."""
    n = len(off) - 1
    e = SgdStressEngine(n_pivots=k, seed=seed)
    e.init(off, nb, np.zeros(3 * n, np.float32))
    piv, dist, _ = e.pivots()
    return StressJudge(n, piv, dist)


class EdgeCv:
    """edge_length_cv (crates/graph-layouts/src/metrics.rs:150-164) over the
    graph's undirected edges (i < j), f64 sums."""

    def __init__(self, off, nb):
        off = np.asarray(off, np.int64)
        src = np.repeat(np.arange(len(off) - 1), np.diff(off))
        keep = src < np.asarray(nb, np.int64)
        self.i, self.j = src[keep], np.asarray(nb, np.int64)[keep]

    def __call__(self, xyz):
        if len(self.i) < 2:
            return 0.0
        p = np.asarray(xyz, np.float32).reshape(-1, 3)
        d = (p[self.i] - p[self.j]).astype(np.float64)
        ln = np.sqrt((d * d).sum(1))
        mean = ln.mean()
        if mean <= 1e-12:
            return 0.0
        return float(np.float32(np.sqrt(((ln - mean) ** 2).mean()) / mean))


def stress_curve(off, nb, init, judge, iters: int, every: int, theta: float,
                 stop_below: float | None = None, cv=None):
    """Runs the plain engine from `init`; stress (and, with `cv`, the edge-length
    CV) after every `every` iterations.  Returns (list of (iteration, stress) or
    (iteration, stress, cv), ms per iteration measured on the device)."""
    e = Fa2BhEngine(theta=theta)
    e.init(off, nb, init)
    curve = [(0, judge(init)) if cv is None else (0, judge(init), cv(init))]
    dev_ms = 0.0
    done = 0
    while done < iters:
        k = min(every, iters - done)
        _, it_ms = e.step_timed(k)
        dev_ms += it_ms
        done += k
        pos = e.get_state()["positions"]
        s = judge(pos)
        curve.append((done, s) if cv is None else (done, s, cv(pos)))
        if stop_below is not None and s <= stop_below:
            break
    return curve, dev_ms / max(done, 1)


def first_within(curve, s_target, cv_band):
    """First check at which stress <= s_target and the edge-length CV lies in
    cv_band = (lo, hi), and whether every later check stays within both."""
    lo, hi = cv_band
    ok = lambda s, c: s <= s_target and lo <= c <= hi  # noqa: E731
    for k, (it, s, c) in enumerate(curve):
        if ok(s, c):
            return it, all(ok(s2, c2) for _, s2, c2 in curve[k:])
    return None, False


def multilevel_curve(off, nb, init, judge, iters: int, every: int, theta: float,
                     stop_below: float | None = None, settings: dict | None = None, cv=None,
                     stop_cv: tuple | None = None):
    """The classical multilevel start (multilevel-hip over fa2-bh-hip, the
    reference's MultilevelSettings defaults unless `settings` overrides them):
    wall time of init (host coarsening, fold-up, first bind) plus the descent
    through every coarse level, then the stress at level 0 after every `every`
    iterations (and, with `cv`, the edge-length CV; the run stops once stress is
    under stop_below and the CV inside the band stop_cv = (lo, hi)).  Returns
    (curve over level-0 iterations, descent wall ms, levels)."""
    st = dict(settings or {})
    e = MultilevelEngine(inner="fa2-bh-hip", inner_params={"theta": theta}, **st)
    t0 = time.perf_counter()
    e.init(off, nb, init)
    levels = e.state()["n_levels"]
    base = int(st.get("sweeps_per_level", 30))
    sched = st.get("sweep_schedule", "linear")
    descent = 0
    for lvl in range(1, levels):
        descent += base if sched == "uniform" else (base * (1 + lvl) if sched == "linear" else base << min(lvl, 16))
    pos = e.step(descent)
    descent_ms = (time.perf_counter() - t0) * 1e3
    assert e.state()["level"] == 0
    curve = [(0, judge(pos)) if cv is None else (0, judge(pos), cv(pos))]
    done = 0
    while done < iters:
        k = min(every, iters - done)
        pos = e.step(k)
        done += k
        s = judge(pos)
        c = None if cv is None else cv(pos)
        curve.append((done, s) if cv is None else (done, s, c))
        if stop_below is not None and s <= stop_below and (stop_cv is None or stop_cv[0] <= c <= stop_cv[1]):
            break
    return curve, descent_ms, levels


def first_below(curve, target):
    for it, s in curve:
        if s <= target:
            return it
    return None


def evaluate(net: Mpnn, coef, off, nb, x0, horizon: int = 2000, every: int = 50,
             theta: float = 0.8, k_pivots: int = 50, ablation: bool = True, multilevel: bool = True,
             log=None):
    """Time to reach stress <= 1.01 x the plain converged stress, plain vs learned,
    and time to reach that together with an edge-length CV within 1 % of the
    plain run's converged CV.

    The plain run starts from x0 and runs `horizon` iterations; its stress there
    is the converged stress S*.  Times are device time of the FA2 iterations
    (HIP events) plus, for the learned run, the whole wall time of the
    prediction (graph upload, features, forward pass, rescale).  The ablation
    starts the plain engine from x0 rescaled to the size law, without the
    network, to separate what the network adds from what the scale adds."""
    judge = stress_judge(off, nb, k_pivots)
    cv = EdgeCv(off, nb)
    plain, ms_it = stress_curve(off, nb, x0, judge, horizon, every, theta, cv=cv)
    s_star, cv_star = plain[-1][1], plain[-1][2]
    target = 1.01 * s_star
    cv_target = (0.99 * cv_star, 1.01 * cv_star)   # within 1 % of the converged CV, either side
    t_plain_it = first_below([c[:2] for c in plain], target)
    t_plain_both, plain_stays = first_within(plain, target, cv_target)
    init, fwd_ms, pred_wall_ms = predict(net, coef, off, nb, x0)
    learned, ms_it2 = stress_curve(off, nb, init, judge, horizon, every, theta, cv=cv)
    t_learn_it = first_below([c[:2] for c in learned], target)
    t_learn_both, learned_stays = first_within(learned, target, cv_target)
    fallback = t_learn_it is None
    res = {
        "horizon": horizon,
        "check_every": every,
        "plain_stress_final": s_star,
        "target_stress": target,
        "plain_init_stress": plain[0][1],
        "plain_iters_to_target": t_plain_it,
        "plain_ms_to_target": t_plain_it * ms_it,
        "learned_init_stress": learned[0][1],
        "learned_stress_final": learned[-1][1],
        "learned_iters_to_target": t_learn_it,
        "inference_ms": pred_wall_ms,
        "inference_forward_device_ms": fwd_ms,
        "learned_ms_to_target": None if fallback else pred_wall_ms + t_learn_it * ms_it2,
        "ms_per_iteration": ms_it,
        "fallback_to_plain": fallback,
        "plain_curve": plain,
        "learned_curve": learned,
        # both judges of the final layout: stress <= 1.01 S* and |CV - CV*| <= 0.01 CV*
        "plain_cv_final": cv_star,
        "target_cv": list(cv_target),
        "plain_init_cv": plain[0][2],
        "learned_init_cv": learned[0][2],
        "learned_cv_final": learned[-1][2],
        "plain_iters_to_both": t_plain_both,
        "plain_ms_to_both": None if t_plain_both is None else t_plain_both * ms_it,
        "plain_stays_within_both": plain_stays,
        "learned_iters_to_both": t_learn_both,
        "learned_ms_to_both": None if t_learn_both is None else pred_wall_ms + t_learn_both * ms_it2,
        "learned_stays_within_both": learned_stays,
        "learned_max_stress_after_target": (None if t_learn_it is None else
                                            max(s2 for it2, s2, _ in learned if it2 >= t_learn_it)),
    }
    res["speedup"] = None if fallback else res["plain_ms_to_target"] / res["learned_ms_to_target"]
    res["speedup_both"] = (None if t_learn_both is None or t_plain_both is None
                           else res["plain_ms_to_both"] / res["learned_ms_to_both"])
    if multilevel:
        # the classical counterpart: coarsen -> solve -> prolong -> refine
        ml, ml_ms, ml_levels = multilevel_curve(off, nb, x0, judge, horizon, every, theta, stop_below=target,
                                                cv=cv, stop_cv=cv_target)
        t_ml = first_below([c[:2] for c in ml], target)
        t_ml_both, _ = first_within(ml, target, cv_target)
        res["multilevel_level0_iters_to_both"] = t_ml_both
        res["multilevel_ms_to_both"] = None if t_ml_both is None else ml_ms + t_ml_both * ms_it
        res["multilevel_levels"] = ml_levels
        res["multilevel_descent_ms"] = ml_ms
        res["multilevel_level0_start_stress"] = ml[0][1]
        res["multilevel_level0_iters_to_target"] = t_ml
        res["multilevel_ms_to_target"] = None if t_ml is None else ml_ms + t_ml * ms_it
        res["multilevel_curve"] = ml
        if log:
            log(f"multilevel: {ml_levels} levels, descent {ml_ms:.0f} ms, level-0 start S={ml[0][1]:.4f}, "
                f"target after {t_ml} more iterations")
    if ablation:
        R = size_of(coef, len(off) - 1, len(nb))
        p = np.asarray(x0, np.float64).reshape(-1, 3)
        xs = ((p - p.mean(0)) * (R / max(rms_radius(p), 1e-30))).astype(np.float32).reshape(-1)
        scaled, ms_it3 = stress_curve(off, nb, xs, judge, horizon, every, theta, cv=cv)
        t_s = first_below([c[:2] for c in scaled], target)
        t_s_both, _ = first_within(scaled, target, cv_target)
        res["scaled_random_iters_to_both"] = t_s_both
        res["scaled_random_ms_to_both"] = None if t_s_both is None else t_s_both * ms_it3
        res["scaled_random_iters_to_target"] = t_s
        res["scaled_random_ms_to_target"] = None if t_s is None else t_s * ms_it3
        res["scaled_random_curve"] = scaled
    if log:
        log(f"plain: S*={s_star:.4f} (init {plain[0][1]:.4f}), CV*={cv_star:.4f} (init {plain[0][2]:.4f}), "
            f"stress target at it {t_plain_it}, both at {t_plain_both}; learned: init S={learned[0][1]:.4f} "
            f"CV={learned[0][2]:.4f}, stress target at it {t_learn_it}, both at {t_learn_both} (stays "
            f"{learned_stays}), final {learned[-1][1]:.4f} / {learned[-1][2]:.4f}, speedup {res['speedup']}, "
            f"both {res['speedup_both']}")
    return res
