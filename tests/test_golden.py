"""The committed golden fixtures (tests/golden/, written by make_golden.py from
the oracle) re-derived on the CPU: integer and f32 results must match bit for
bit, the float64 network output to rounding.  A change in the oracle, the
generators or the fixture data fails here before any GPU comparison uses it.
The product's host-side generators and CSR encoding are checked against the
same digests."""
import importlib.util
import json
import os

import numpy as np
import pytest

import jumpcannon as jc
from oracle import mpnn_oracle as M

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
G = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(G)


def _load(name):
    with np.load(os.path.join(HERE, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _bits_equal(a, b):
    a = np.ascontiguousarray(np.asarray(a)).reshape(-1)
    b = np.ascontiguousarray(np.asarray(b)).reshape(-1)
    assert a.shape == b.shape and a.dtype == b.dtype
    return a.tobytes() == b.tobytes()


@pytest.mark.parametrize("name", ["fa2_sbm2000.npz", "octree_cube1000.npz",
                                  "pagerank_ba3000.npz", "sgd_rgg2000.npz"])
def test_oracle_reproduces_fixture(name):
    want = _load(name)
    got = G.FIXTURES[name]()
    assert set(want) == set(got)
    for k, v in got.items():
        w = want[k]
        if w.dtype.kind in "U":
            assert str(w) == v, k
        else:
            assert _bits_equal(np.asarray(v, dtype=w.dtype), w), k


def test_mpnn_oracle_reproduces_fixture():
    want = _load("mpnn_tiny.npz")
    off, nb, params, x, tgt = G.mpnn_inputs()
    assert len(params) == int(want["n_params"])
    y, _ = M.forward(params, off, nb, x, emulate_bf16=False)
    L, g = M.loss_grad(params, off, nb, x, tgt, [0, 7], emulate_bf16=False)
    np.testing.assert_allclose(y, want["y"], rtol=1e-12, atol=1e-14)
    assert abs(L - float(want["loss"])) <= 1e-12 * abs(L)
    np.testing.assert_allclose(g[::G.GRAD_STRIDE], want["grad_sample"], rtol=1e-10, atol=1e-14)


def test_fixture_facts():
    """Independent facts about the fixtures: the octree holds every body once and
    its leaf masses add up to sum(deg + 1); ranks sum to 1 within f32 rounding;
    the SGD run kept z = 0."""
    t = _load("octree_cube1000.npz")
    leaves = t["body"] != 0xFFFFFFFF
    assert np.array_equal(np.sort(t["body"][leaves]), np.arange(1000, dtype=np.uint32))
    assert np.array_equal(np.sort(t["perm"]), np.arange(1000, dtype=np.uint32))
    assert np.all(np.diff(t["keys"].astype(np.uint64)) >= 0)
    root_mass = float(t["node_a"][0, 3])
    assert abs(root_mass - float(t["node_a"][leaves, 3].astype(np.float64).sum())) < 1e-3 * root_mass
    pr = _load("pagerank_ba3000.npz")
    assert abs(pr["ranks"].astype(np.float64).sum() - 1.0) < 1e-4 and pr["residual"] < 1e-6
    sg = _load("sgd_rgg2000.npz")
    assert np.abs(sg["pos10"][2::3]).max() == 0.0


def test_product_generators_and_csr_encoding_match_digests():
    gen = json.load(open(os.path.join(HERE, "generators.json")))
    off, nb = jc.sbm_graph(2000, 4, 0.02, 0.001, 0xC0FFEE)
    assert G.digest(off, nb) == gen["sbm_2000_4"]["sha256"]
    assert G.digest(jc.uniform_cube(1_000_000, 10.0, 0xC0FFEE + 1)) == gen["cube_1000000_10_seed+1"]
    u, h = jc._splitmix_pair(0)
    first = gen["splitmix_0_first4"]
    assert h == first[1] and u == (first[0] >> 11) / float(1 << 53)
    fa2 = _load("fa2_sbm2000.npz")
    assert G.digest(off, nb) == str(fa2["graph_sha256"])
    o2, n2 = jc.csr_from_bin(jc.csr_to_bin(off, nb))
    assert G.digest(o2, n2) == gen["sbm_2000_4"]["sha256"]


def test_product_ba_1m_matches_digest():
    gen = json.load(open(os.path.join(HERE, "generators.json")))["ba_1000000_4"]
    off, nb = jc.ba_graph(1_000_000, 4, 0xC0FFEE)
    assert len(nb) == gen["nnz"] and G.digest(off, nb) == gen["sha256"]
