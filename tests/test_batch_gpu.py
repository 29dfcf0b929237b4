"""Batched FA2+BH: every graph of a batch must evolve bit-identically to the same
graph run alone through fa2-bh-hip (and hence to the oracle), whatever the mix
of sizes, including single-body, empty-edge and non-finite cases."""
import numpy as np
import pytest

import jumpcannon as jc
from jumpcannon import workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _mixed(n_graphs, seed, max_n=3000):
    rng = np.random.default_rng(seed)
    graphs, pos = [], []
    for g in range(n_graphs):
        n = int(rng.integers(1, max_n))
        if g % 3 == 0:
            off, nb = O.sbm_graph(n, 3, 0.05, 0.002, seed + g)
        elif g % 3 == 1:
            off, nb = O.ba_graph(max(n, 2), 2 + g % 3, seed + g)
        else:
            off, nb = O.csr_from_pairs(n, np.zeros(0, np.uint32))  # isolated bodies
        graphs.append((off, nb))
        pos.append(O.uniform_cube(len(off) - 1, 1.0 + g % 5, seed + 1000 + g))
    return graphs, pos


@pytest.mark.parametrize("seed", [1, 2])
def test_batch_equals_standalone_runs(seed):
    graphs, pos = _mixed(24, seed)
    pos[5] = pos[5].copy()
    pos[5][4] = np.nan  # a non-finite body inside one graph
    b = jc.Fa2BhBatch(theta=1.0)
    b.init(graphs, pos)
    out = b.step(25)
    se = b.speeds()
    for g, ((off, nb), p0) in enumerate(zip(graphs, pos)):
        e = jc.Fa2BhEngine(theta=1.0)
        e.init(off, nb, p0)
        want = e.step(25)
        assert np.array_equal(out[g].view(np.uint32), want.view(np.uint32)), f"graph {g}"
        st = e.get_state()
        assert np.array_equal(se[g].view(np.uint32), st["speed_eff2"].view(np.uint32))


def test_batch_matches_oracle_trajectories():
    graphs, pos = _mixed(8, 7, max_n=800)
    b = jc.Fa2BhBatch(theta=1.0)
    b.init(graphs, pos)
    out = b.step(12)
    for g, ((off, nb), p0) in enumerate(zip(graphs, pos)):
        want, _, _, _ = O.fa2_run(off, nb, p0, 12, O.Fa2Settings.make(theta=1.0), tree_kind=1,
                                  exact=True)
        assert np.array_equal(out[g].view(np.uint32), want.view(np.uint32)), f"graph {g}"


def test_batch_single_graph_and_reuse():
    off, nb = O.ba_graph(5000, 3, 4)
    p0 = O.uniform_cube(5000, 2.0, 5)
    b = jc.Fa2BhBatch(theta=0.9)
    b.init([(off, nb)], [p0])
    a = b.step(7)[0]
    e = jc.Fa2BhEngine(theta=0.9)
    e.init(off, nb, p0)
    assert np.array_equal(a.view(np.uint32), e.step(7).view(np.uint32))


def test_batch_4096_sampled_graphs_equal_standalone():
    """The 4,096-graph workload (log-uniform 50..20,000, half SBM, half BA) for
    its configured 500 iterations: run the whole batch, then check a spread of
    graphs (smallest, largest, both ends, the middle, six seeded picks) against
    standalone runs of the same 500 iterations, bit for bit."""
    graphs, pos = workloads.batch_4096()
    b = jc.Fa2BhBatch(theta=1.0)
    b.init(graphs, pos)
    out = b.step(500)
    sizes = np.array([len(o) - 1 for o, _ in graphs])
    pick = [int(np.argmin(sizes)), int(np.argmax(sizes)), 0, 1, 2047, 4095]
    pick += [int(g) for g in np.random.default_rng(4096).choice(4096, 6, replace=False)]
    for g in pick:
        off, nb = graphs[g]
        e = jc.Fa2BhEngine(theta=1.0)
        e.init(off, nb, pos[g])
        assert np.array_equal(out[g].view(np.uint32), e.step(500).view(np.uint32)), f"graph {g}"
    assert all(np.isfinite(x).all() for x in out)
