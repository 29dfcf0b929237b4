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


# (schedule, persistent workgroup size): both schedules must give the same bits.
# The persistent schedule is switched off in the library (its first GPU run made
# a memory fault whose cause has not been found), so its cases are skipped.
PERSISTENT_OFF = "persistent schedule switched off: its first GPU run faulted, cause not found"
_P = pytest.mark.skip(reason=PERSISTENT_OFF)
SCHEDULES = [("phased", 0), pytest.param(("persistent", 256), marks=_P),
             pytest.param(("persistent", 512), marks=_P), pytest.param(("persistent", 1024), marks=_P)]
SCHED_NAMES = ["phased", pytest.param("persistent", marks=_P)]


@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: f"{s[0]}{s[1] or ''}")
@pytest.mark.parametrize("seed", [1, 2])
def test_batch_equals_standalone_runs(seed, sched):
    graphs, pos = _mixed(24, seed)
    pos[5] = pos[5].copy()
    pos[5][4] = np.nan  # a non-finite body inside one graph
    b = jc.Fa2BhBatch(theta=1.0)
    b.set_schedule(*sched)
    b.init(graphs, pos)
    out = b.step(25)
    assert b.schedule()[:2] == sched if sched[0] == "persistent" else b.schedule()[0] == "phased"
    se = b.speeds()
    for g, ((off, nb), p0) in enumerate(zip(graphs, pos)):
        e = jc.Fa2BhEngine(theta=1.0)
        e.init(off, nb, p0)
        want = e.step(25)
        assert np.array_equal(out[g].view(np.uint32), want.view(np.uint32)), f"graph {g}"
        st = e.get_state()
        assert np.array_equal(se[g].view(np.uint32), st["speed_eff2"].view(np.uint32))


@pytest.mark.parametrize("sched", SCHED_NAMES)
def test_batch_matches_oracle_trajectories(sched):
    graphs, pos = _mixed(8, 7, max_n=800)
    b = jc.Fa2BhBatch(theta=1.0)
    b.set_schedule(sched)
    b.init(graphs, pos)
    out = b.step(12)
    for g, ((off, nb), p0) in enumerate(zip(graphs, pos)):
        want, _, _, _ = O.fa2_run(off, nb, p0, 12, O.Fa2Settings.make(theta=1.0), tree_kind=1,
                                  exact=True)
        assert np.array_equal(out[g].view(np.uint32), want.view(np.uint32)), f"graph {g}"


@pytest.mark.parametrize("sched", SCHED_NAMES)
def test_batch_single_graph_and_reuse(sched):
    off, nb = O.ba_graph(5000, 3, 4)
    p0 = O.uniform_cube(5000, 2.0, 5)
    b = jc.Fa2BhBatch(theta=0.9)
    b.set_schedule(sched)
    b.init([(off, nb)], [p0])
    a = b.step(7)[0]
    e = jc.Fa2BhEngine(theta=0.9)
    e.init(off, nb, p0)
    assert np.array_equal(a.view(np.uint32), e.step(7).view(np.uint32))
    # the same batch continues: 3 single steps then 4 in one call
    for _ in range(3):
        a = b.step(1)[0]
    a = b.step(4)[0]
    assert np.array_equal(a.view(np.uint32), e.step(7).view(np.uint32))


@pytest.mark.skip(reason=PERSISTENT_OFF)
def test_batch_schedules_switch_mid_run():
    """Phased and persistent iterations interleaved on one batch give the bits of
    one standalone run (the schedules share the positions, forces and speeds)."""
    graphs, pos = _mixed(10, 11, max_n=1500)
    b = jc.Fa2BhBatch(theta=1.0)
    b.init(graphs, pos)
    for i, sched in enumerate(["persistent", "phased", "persistent", "auto"]):
        b.set_schedule(sched)
        out = b.step(3 + i)
    for g, ((off, nb), p0) in enumerate(zip(graphs, pos)):
        e = jc.Fa2BhEngine(theta=1.0)
        e.init(off, nb, p0)
        assert np.array_equal(out[g].view(np.uint32), e.step(18).view(np.uint32)), f"graph {g}"


@pytest.mark.skip(reason=PERSISTENT_OFF)
def test_batch_persistent_edge_shapes():
    """Single-body graphs, an empty-edge graph, coincident bodies (equal keys),
    a hub with rows longer than the short-row limit, and all bodies non-finite."""
    graphs, pos = [], []
    off, nb = O.csr_from_pairs(1, np.zeros(0, np.uint32))
    graphs.append((off, nb))
    pos.append(np.array([0.5, -0.25, 2.0], np.float32))
    hub = np.array([[0, j] for j in range(1, 700)] + [[j, j + 1] for j in range(1, 699)], np.uint32)
    graphs.append(O.csr_from_pairs(700, hub.reshape(-1)))
    pos.append(O.uniform_cube(700, 3.0, 21))
    dup = O.uniform_cube(300, 1.0, 22).reshape(-1, 3)
    dup[100:200] = dup[100]  # 100 coincident bodies
    graphs.append(O.ba_graph(300, 2, 23))
    pos.append(dup.reshape(-1))
    graphs.append(O.csr_from_pairs(5, np.zeros(0, np.uint32)))
    pos.append(np.full(15, np.nan, np.float32))
    graphs.append(O.sbm_graph(2500, 4, 0.01, 0.001, 24))
    pos.append(O.uniform_cube(2500, 10.0, 25))
    b = jc.Fa2BhBatch(theta=1.0)
    b.set_schedule("persistent", 256)
    b.init(graphs, pos)
    out = b.step(9)
    se = b.speeds()
    for g, ((off, nb), p0) in enumerate(zip(graphs, pos)):
        e = jc.Fa2BhEngine(theta=1.0)
        e.init(off, nb, p0)
        want = e.step(9)
        assert np.array_equal(out[g].view(np.uint32), want.view(np.uint32)), f"graph {g}"
        assert np.array_equal(se[g].view(np.uint32), e.get_state()["speed_eff2"].view(np.uint32))


@pytest.fixture(scope="module")
def knob_case():
    """A mixed batch and its graphs' standalone trajectories (10 iterations)."""
    graphs, pos = _mixed(14, 31, max_n=2500)
    want = []
    for (off, nb), p0 in zip(graphs, pos):
        e = jc.Fa2BhEngine(theta=1.0)
        e.init(off, nb, p0)
        want.append(e.step(10))
    return graphs, pos, want


# (force block, centre-of-mass fusion depth, chunk, sort key width 0 = narrowest / 1 = +16)
KNOBS = [(64, 3, 1024, 0), (1024, 3, 2048, 1), (256, 4, 1024, 1), (512, 4, 2048, 0),
         (128, 5, 1024, 0), (1024, 5, 2048, 1)]


@pytest.mark.parametrize("knobs", KNOBS, ids=lambda k: "fb{}-d{}-c{}-w{}".format(*k))
def test_batch_every_tuning_knob_gives_the_same_bits(knob_case, knobs):
    """Each setting the autotuner can pick is bit-identical to the standalone runs."""
    graphs, pos, want = knob_case
    fb, depth, chunk, wide = knobs
    b = jc.Fa2BhBatch(theta=1.0)
    b.init(graphs, pos)
    bits = b.schedule()[5]  # the narrowest key width this batch allows
    b2 = jc.Fa2BhBatch(theta=1.0)
    b2.set_tuning(force_block=fb, com_depth=depth, com_chunk=chunk,
                  sort_key_bits=min(bits + 16 * wide, 64))
    b2.init(graphs, pos)
    out = b2.step(10)
    sched = b2.schedule()
    assert sched[2:5] == (fb, depth, chunk)
    assert sched[5] == min(bits + 16 * wide, 64)
    for g in range(len(graphs)):
        assert np.array_equal(out[g].view(np.uint32), want[g].view(np.uint32)), f"graph {g}"


def test_batch_4096_sampled_graphs_equal_standalone():
    """The 4,096-graph workload (log-uniform 50..20,000, half SBM, half BA) for
    its configured 500 iterations: run the whole batch, then check a spread of
    graphs (smallest, largest, both ends, the middle, six seeded picks) against
    standalone runs of the same 500 iterations, bit for bit."""
    graphs, pos = workloads.batch_4096()
    b = jc.Fa2BhBatch(theta=1.0)
    b.init(graphs, pos)
    out = b.step(500)
    print("schedule", b.schedule())
    sizes = np.array([len(o) - 1 for o, _ in graphs])
    pick = [int(np.argmin(sizes)), int(np.argmax(sizes)), 0, 1, 2047, 4095]
    pick += [int(g) for g in np.random.default_rng(4096).choice(4096, 6, replace=False)]
    for g in pick:
        off, nb = graphs[g]
        e = jc.Fa2BhEngine(theta=1.0)
        e.init(off, nb, pos[g])
        assert np.array_equal(out[g].view(np.uint32), e.step(500).view(np.uint32)), f"graph {g}"
    assert all(np.isfinite(x).all() for x in out)
