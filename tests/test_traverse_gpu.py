"""BFS, connected components and the largest component on the GPU, bit-exact
against the oracle (queue BFS of analytics/bfs.rs:44-66, union-find of
analytics/connected_components.rs:45-66), on the shapes of the reference's own
analytics tests and at the workload sizes: the 1M-node scale-free graph
(frontiers of hundreds of thousands: whole-chip levels) and the 200k-node
random geometric graph (hundreds of thin levels: the one-workgroup kernel)."""
import numpy as np
import pytest

import jumpcannon as jc
from jumpcannon import workloads
from oracle import oracle as O
import traverse_graphs as G

pytestmark = pytest.mark.gpu

UNREACHABLE = 0xFFFFFFFF


def _bfs_eq(off, nb, src):
    got = jc.bfs(off, nb, src)
    exp = O.bfs(off, nb, src)
    assert got.dtype == np.uint32 and np.array_equal(got, exp), np.nonzero(got != exp)[0][:10]
    return got


def test_bfs_reference_shapes():
    """gpu_bfs.rs: path from 10, disconnected node, grid from the corner."""
    off, nb = G.path(50)
    assert _bfs_eq(off, nb, 10).tolist() == [abs(v - 10) for v in range(50)]
    off = np.array([0, 1, 3, 4, 4], np.uint32)
    nb = np.array([1, 0, 2, 1], np.uint32)
    assert _bfs_eq(off, nb, 0).tolist() == [0, 1, 2, UNREACHABLE]
    off, nb = G.grid(5, 6)
    assert _bfs_eq(off, nb, 0).tolist() == [r + c for r in range(5) for c in range(6)]


@pytest.mark.parametrize("seed", range(24))
def test_bfs_random_graphs(seed):
    """gpu_bfs.rs gpu_equals_cpu_bfs: random graphs from a random source."""
    (off, nb), src = G.random_graph(seed)
    _bfs_eq(off, nb, src)


def test_bfs_errors_and_empty():
    off, nb = G.path(5)
    with pytest.raises(jc.JcError) as e:
        jc.bfs(off, nb, 5)
    assert e.value.code == jc.JC_E_INVALID and "out of range" in str(e.value)
    assert jc.bfs(np.zeros(1, np.uint32), np.zeros(0, np.uint32), 0).size == 0


def test_bfs_frontier_spills_and_returns():
    """Hub rows longer than a lane's share, a frontier beyond the LDS queue
    (whole-chip level), then back to one node (the workgroup kernel rebuilds its
    visited bitmap from the distances)."""
    off, nb = G.broom(6000, 100)
    for src in (0, 1, 6001, len(off) - 2):
        _bfs_eq(off, nb, src)
    off, nb = G.star(20000)
    _bfs_eq(off, nb, 0)
    _bfs_eq(off, nb, 777)


def test_bfs_directed_hub_stays_top_down():
    """A directed CSR whose 10,000-node frontier would go bottom-up on a
    symmetric graph: distances follow out-edges only, as cpu_bfs does
    (bfs.rs:44-66).  Also the symmetric star with its rows shuffled."""
    off, nb = G.directed_hub(10000)
    d = _bfs_eq(off, nb, 0)
    assert d[10001:15001].tolist() == [2] * 5000
    assert (d[15001:] == UNREACHABLE).all()
    _bfs_eq(*G.shuffled_rows(*G.star(20000)), 0)
    _bfs_eq(*G.shuffled_rows(*G.broom(6000, 100)), 1)


def test_bfs_ba_1m():
    """The 1M-node scale-free workload graph (above the LDS bitmap size, so the
    workgroup kernel claims nodes on the distance array), several sources."""
    off, nb, _ = workloads.ba_1m()
    plan = jc.BfsPlan(off, nb)
    for src in (0, 123_457, 999_999):
        d, ms = plan.run(src)
        exp = O.bfs(off, nb, src)
        assert np.array_equal(d, exp)
        print(f"BFS BA-1M from {src}: {ms:.3f} ms, max hop {int(exp.max())}")


def test_bfs_rgg_200k():
    off, nb, _ = workloads.rgg_200k()
    plan = jc.BfsPlan(off, nb)
    for src in (0, 77_777):
        d, ms = plan.run(src)
        exp = O.bfs(off, nb, src)
        assert np.array_equal(d, exp)
        print(f"BFS RGG-200k from {src}: {ms:.3f} ms, {int(exp.max())} levels")


def _cc_eq(off, nb):
    got = jc.connected_components(off, nb)
    exp = O.connected_components(off, nb)
    assert np.array_equal(got, exp), np.nonzero(got != exp)[0][:10]
    return got


def test_cc_reference_shapes():
    """gpu_connected_components.rs: cliques, path, isolated nodes, two cycles."""
    assert _cc_eq(*G.cliques(5, 7)).tolist() == [7 * (v // 7) for v in range(35)]
    assert _cc_eq(*G.path(64)).tolist() == [0] * 64
    assert _cc_eq(np.zeros(4, np.uint32), np.zeros(0, np.uint32)).tolist() == [0, 1, 2]
    assert len(set(_cc_eq(*G.cycles(10, 12)).tolist())) == 2
    _cc_eq(*G.directed_only())
    assert jc.connected_components(np.zeros(1, np.uint32), np.zeros(0, np.uint32)).size == 0


@pytest.mark.parametrize("seed", range(24))
def test_cc_random_graphs(seed):
    (off, nb), _ = G.random_graph(seed)
    _cc_eq(off, nb)


def test_cc_large():
    off, nb, _ = workloads.ba_1m()
    assert (_cc_eq(off, nb) == 0).all()
    # many components of every size: a sparse SBM and its hubs-free remainder
    off, nb = O.sbm_graph(200_000, 50, 3.0 / 4000, 0.2 / 200_000, 17)
    lab = _cc_eq(off, nb)
    print(f"SBM 200k: {len(np.unique(lab))} components")


def test_largest_component_matches_oracle_and_workload():
    off, nb = O.sbm_graph(50_000, 10, 2.5 / 5000, 0.3 / 50_000, 3)
    keep, so, sn = jc.largest_component(off, nb)
    size, okeep = O.largest_component(len(off) - 1, off, nb)
    assert np.array_equal(keep, okeep) and len(so) == size + 1
    # the induced CSR equals the relabelled pairs rebuilt on the host
    sel = okeep != UNREACHABLE
    rows = np.repeat(np.arange(len(off) - 1), np.diff(off))
    m = sel[rows]
    pairs = np.stack([okeep[rows[m]], okeep[nb[m]]], 1).reshape(-1)
    eo, en = O.csr_from_pairs(size, pairs)
    assert np.array_equal(so, eo) and np.array_equal(sn, en)
    # the stress workload's largest component, built through scipy on the host
    xy, pairs = jc.rgg_pairs(200_000, np.sqrt(8.0 / (np.pi * 200_000)), workloads.SEED)
    roff, rnb = jc.csr_from_pairs(200_000, pairs)
    w_off, w_nb, _ = workloads.rgg_200k()
    _, so, sn = jc.largest_component(roff, rnb)
    assert np.array_equal(so, w_off) and np.array_equal(sn, w_nb)


def test_largest_component_ties_and_empty():
    keep, so, sn = jc.largest_component(*G.cycles(8, 8))
    assert keep.tolist() == list(range(8)) + [UNREACHABLE] * 8
    assert so.tolist() == G.cycles(8)[0].tolist()
    keep, so, sn = jc.largest_component(np.zeros(1, np.uint32), np.zeros(0, np.uint32))
    assert keep.size == 0 and so.tolist() == [0] and sn.size == 0


def test_sgd_pivots_on_disconnected_graph():
    """Max-min pivots over several components and isolated nodes: unreachable
    rows are 0 and never raise a node's cover (sgd_stress.rs:143-175)."""
    off, nb = G.cycles(10, 12, 3)
    off = np.concatenate([off, np.full(4, off[-1], np.uint32)])  # 4 isolated nodes
    n = len(off) - 1
    for k in (1, 5, n):
        e = jc.SgdStressEngine(n_pivots=k)
        e.init(off, nb, O.uniform_cube(n, 1.0, 2))
        piv, dist, _ = e.pivots()
        opiv, odist = O.select_pivots(off, nb, k)
        assert np.array_equal(piv, opiv) and np.array_equal(dist, odist)


def test_sgd_pivot_distance_beyond_u16_is_rejected():
    """A path of 140,000 nodes has a hop distance above 65535 from every node."""
    off, nb = O.csr_from_pairs(140_000, np.stack([np.arange(139_999), np.arange(1, 140_000)], 1)
                               .astype(np.uint32).reshape(-1))
    e = jc.SgdStressEngine(n_pivots=1)
    with pytest.raises(jc.JcError) as err:
        e.init(off, nb, np.zeros(3 * 140_000, np.float32))
    assert "u16" in str(err.value)


def test_bfs_large_grid_many_levels_over_the_csr():
    """An 800 x 700 grid (560,000 nodes, above the LDS bitmap's 524,288) keeps
    its frontier under the LDS queue for ~1,500 levels: the workgroup loop over
    the CSR with atomicCAS claims.  From a corner the distance is Manhattan."""
    rows, cols = 700, 800
    v = np.arange(rows * cols, dtype=np.uint32).reshape(rows, cols)
    pairs = np.concatenate([np.stack([v[:, :-1].ravel(), v[:, 1:].ravel()], 1),
                            np.stack([v[:-1, :].ravel(), v[1:, :].ravel()], 1)]).astype(np.uint32)
    off, nb = O.csr_from_pairs(rows * cols, pairs.reshape(-1))
    d = jc.bfs(off, nb, 0)
    r, c = np.divmod(np.arange(rows * cols), cols)
    assert np.array_equal(d, (r + c).astype(np.uint32))
    mid = (rows // 2) * cols + cols // 2
    assert np.array_equal(jc.bfs(off, nb, mid), O.bfs(off, nb, mid))
    assert (jc.connected_components(off, nb) == 0).all()
