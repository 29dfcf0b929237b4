"""The GPU path against the committed golden fixtures (tests/golden/, written
from the oracle by make_golden.py): forces, exact sums and trajectories of C1,
an octree, PageRank and an SGD run, all bit for bit.  Inputs are regenerated
by the library's own host generators and checked against the fixtures'
digests first."""
import importlib.util
import os

import numpy as np
import pytest

import jumpcannon as jc
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
G = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(G)


def _load(name):
    with np.load(os.path.join(HERE, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _same(a, b):
    return np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()


def test_fa2_sbm2000_forces_sums_and_trajectory():
    f = _load("fa2_sbm2000.npz")
    off, nb = jc.sbm_graph(2000, 4, 0.02, 0.001, 0xC0FFEE)
    x0 = jc.uniform_cube(2000, 1.0, 0xC0FFEE + 1)
    assert G.digest(off, nb) == str(f["graph_sha256"]) and G.digest(x0) == str(f["x0_sha256"])
    e = jc.Fa2BhEngine(theta=1.2)
    e.init(off, nb, x0)
    forces, sums = e.forces_at(x0)
    assert _same(forces, f["forces0"]) and _same(sums, f["sums0"])
    e.step(5, readback=False)
    st = e.get_state()
    assert _same(st["positions"], f["pos5"]) and _same(st["old_force"], f["old5"])
    assert _same(st["speed_eff2"], f["se5"])
    e.step(45, readback=False)
    st = e.get_state()
    assert G.digest(st["positions"]) == str(f["pos50_sha256"])
    assert G.digest(st["old_force"]) == str(f["old50_sha256"]) and _same(st["speed_eff2"], f["se50"])


def test_octree_cube1000():
    t = _load("octree_cube1000.npz")
    off, nb = jc.csr_from_pairs(1000, jc.ba_pairs(1000, 2, 1007))
    xyz = jc.uniform_cube(1000, 3.0, 1011)
    assert G.digest(off, nb) == str(t["graph_sha256"]) and G.digest(xyz) == str(t["x_sha256"])
    e = jc.Fa2BhEngine(theta=0.8)
    e.init(off, nb, xyz)
    e.forces_at(xyz)
    g = e.tree()
    N = int(t["n_tree"])
    assert g["n_tree"] == N
    assert np.array_equal(g["keys"], t["keys"]) and np.array_equal(g["perm"], t["perm"])
    assert _same(g["node_a"], t["node_a"])
    internal = t["body"] == O.INTERNAL
    h2 = (t["half"] * np.float32(2.0)) * (t["half"] * np.float32(2.0))
    want_info = np.where(internal, h2.astype(np.float32).view(np.uint32), np.uint32(0x80000000) | t["body"])
    assert np.array_equal(g["info"], want_info.astype(np.uint32))
    assert np.array_equal(g["skip"], t["skip"])
    assert np.array_equal(g["depth"], np.where(internal, t["depth"], 0x80 | t["depth"]).astype(np.uint8))


def test_pagerank_ba3000():
    p = _load("pagerank_ba3000.npz")
    off, nb = jc.ba_graph(3000, 4, 0xC0FFEE)
    assert G.digest(off, nb) == str(p["graph_sha256"])
    ranks, iters, resid = jc.pagerank(off, nb, 0.85, 100, 1e-6)
    assert _same(ranks, p["ranks"]) and iters == int(p["iters"])
    assert np.float64(resid).tobytes() == p["residual"].tobytes()


def test_sgd_rgg2000():
    s = _load("sgd_rgg2000.npz")
    e = jc.SgdStressEngine(n_pivots=20, eta_max=0.0, eta_min=0.01, n_anneal_steps=10)
    e.init(s["offsets"], s["neighbors"], s["x0"])
    piv, dist, _ = e.pivots()
    assert np.array_equal(piv, s["pivots"]) and G.digest(dist) == str(s["dist_sha256"])
    assert _same(e.step(10), s["pos10"])
