"""Pins the oracle's BFS, connected components and largest component before
they check the GPU: the closed forms of the reference's analytics tests
(crates/graph-compute/tests/gpu_bfs.rs, gpu_connected_components.rs) and
scipy's independent graph routines on random graphs."""
import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components, shortest_path

from oracle import oracle as O
import traverse_graphs as G

UNREACHABLE = 0xFFFFFFFF


def _scipy_bfs(off, nb, src):
    n = len(off) - 1
    a = csr_matrix((np.ones(len(nb)), nb.astype(np.int64), off.astype(np.int64)), shape=(n, n))
    d = shortest_path(a, unweighted=True, directed=True, indices=[src])[0]
    out = np.full(n, UNREACHABLE, np.uint32)
    fin = np.isfinite(d)
    out[fin] = d[fin].astype(np.uint32)
    return out


def _scipy_min_labels(off, nb):
    n = len(off) - 1
    a = csr_matrix((np.ones(len(nb)), nb.astype(np.int64), off.astype(np.int64)), shape=(n, n))
    _, lab = connected_components(a, directed=False)
    first = np.full(lab.max() + 1 if n else 0, n, np.int64)
    np.minimum.at(first, lab, np.arange(n))
    return first[lab].astype(np.uint32)


def test_bfs_path_is_linear():
    """gpu_bfs.rs path_distances_are_linear: path of 50 from node 10."""
    off, nb = G.path(50)
    assert O.bfs(off, nb, 10).tolist() == [abs(v - 10) for v in range(50)]


def test_bfs_disconnected_node_unreachable():
    """gpu_bfs.rs disconnected_node_is_unreachable."""
    off = np.array([0, 1, 3, 4, 4], np.uint32)
    nb = np.array([1, 0, 2, 1], np.uint32)
    assert O.bfs(off, nb, 0).tolist() == [0, 1, 2, UNREACHABLE]


def test_bfs_grid_is_manhattan_and_star():
    off, nb = G.grid(5, 6)
    d = O.bfs(off, nb, 0)
    assert d.tolist() == [r + c for r in range(5) for c in range(6)]
    off, nb = G.star(9)
    assert O.bfs(off, nb, 3).tolist() == [1] + [2] * 2 + [0] + [2] * 6


def test_bfs_source_out_of_range_reaches_nothing():
    """bfs.rs:48-50."""
    off, nb = G.path(4)
    assert O.bfs(off, nb, 9).tolist() == [UNREACHABLE] * 4


@pytest.mark.parametrize("seed", range(24))
def test_bfs_matches_scipy_on_random_graphs(seed):
    (off, nb), src = G.random_graph(seed)
    assert np.array_equal(O.bfs(off, nb, src), _scipy_bfs(off, nb, src))


def test_bfs_broom_and_ba_match_scipy():
    off, nb = G.broom(300, 20)
    for src in (0, len(off) - 2):
        assert np.array_equal(O.bfs(off, nb, src), _scipy_bfs(off, nb, src))
    off, nb = O.ba_graph(5000, 3, 21)
    assert np.array_equal(O.bfs(off, nb, 17), _scipy_bfs(off, nb, 17))


def test_cc_closed_forms():
    """gpu_connected_components.rs: 5 cliques of 7 labelled 0,7,..,28; a path is
    one component labelled 0; isolated nodes label themselves; two cycles."""
    off, nb = G.cliques(5, 7)
    assert O.connected_components(off, nb).tolist() == [7 * (v // 7) for v in range(35)]
    off, nb = G.path(64)
    assert O.connected_components(off, nb).tolist() == [0] * 64
    off = np.zeros(4, np.uint32)
    assert O.connected_components(off, np.zeros(0, np.uint32)).tolist() == [0, 1, 2]
    off, nb = G.cycles(10, 12)
    assert O.connected_components(off, nb).tolist() == [0] * 10 + [10] * 12


@pytest.mark.parametrize("seed", range(24))
def test_cc_matches_scipy_on_random_graphs(seed):
    (off, nb), _ = G.random_graph(seed)
    assert np.array_equal(O.connected_components(off, nb), _scipy_min_labels(off, nb))


def test_cc_non_symmetric_rows_are_undirected():
    off, nb = G.directed_only()
    assert np.array_equal(O.connected_components(off, nb), _scipy_min_labels(off, nb))


def test_largest_component_matches_scipy():
    off, nb = O.sbm_graph(3000, 6, 0.004, 0.0001, 9)
    size, keep = O.largest_component(len(off) - 1, off, nb)
    lab = _scipy_min_labels(off, nb)
    ids, counts = np.unique(lab, return_counts=True)
    best = ids[np.lexsort((ids, -counts))[0]]  # largest, then smallest label
    sel = lab == best
    assert size == int(sel.sum())
    exp = np.full(len(lab), UNREACHABLE, np.uint32)
    exp[sel] = np.arange(size, dtype=np.uint32)
    assert np.array_equal(keep, exp)
    # ties: two equal cycles -> the one holding node 0
    off, nb = G.cycles(8, 8)
    size, keep = O.largest_component(16, off, nb)
    assert size == 8 and keep.tolist() == list(range(8)) + [UNREACHABLE] * 8


def test_bfs_directed_hub_follows_out_edges():
    """The directed hub of the GPU's top-down-only case: the oracle follows
    out-edges (bfs.rs:44-66), as scipy's directed search does."""
    off, nb = G.directed_hub(10000)
    d = O.bfs(off, nb, 0)
    assert np.array_equal(d, _scipy_bfs(off, nb, 0))
    assert d[1:10001].tolist() == [1] * 10000 and d[10001:15001].tolist() == [2] * 5000
    assert (d[15001:] == UNREACHABLE).all()
    so, sn = G.shuffled_rows(*G.star(300))
    assert np.array_equal(O.bfs(so, sn, 0), _scipy_bfs(so, sn, 0))
