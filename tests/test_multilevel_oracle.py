"""Pins the oracle's restatement of the multilevel cascade
(crates/graph-layouts/src/layout/coarsen.rs, crates/graph-compute/src/engines/
multilevel.rs) before it checks the GPU wrapper: the reference's own unit tests,
hand-worked matchings, the cascade's invariants on a larger graph, and the
jitter stream against an independent big-integer restatement of xorshift64*."""
import numpy as np
import pytest

from oracle import oracle as O
from jumpcannon import workloads


def test_reference_unit_tests():
    # coarsen.rs coarsen_collapses_a_path
    lv = O.coarsen(6, [0, 1, 1, 2, 2, 3, 3, 4, 4, 5], 4, 2)
    assert lv[0]["n"] == 6 and len(lv) >= 2
    assert len(lv[1]["parent_map"]) == 6 and lv[1]["n"] < 6
    # coarsen.rs prolong_inherits_parent_position
    assert O.prolong([10.0, 20.0, 30.0], [0, 0], 2, 0.0, 7).tolist() == [10.0, 20.0, 30.0] * 2
    # multilevel.rs level_csr_is_symmetric
    off, nb = O.level_csr(3, [0, 1, 1, 2])
    assert off.tolist() == [0, 1, 3, 4] and sorted(nb[1:3].tolist()) == [0, 2]


def test_greedy_matching_by_hand():
    """Path 0-1-2-3-4-5 in edge order: (0,1), (2,3), (4,5) match; the contracted
    edges come from (1,2) and (3,4) in that order."""
    ns, pm, e = O.contract_level(6, [0, 1, 1, 2, 2, 3, 3, 4, 4, 5])
    assert ns == 3 and pm.tolist() == [0, 0, 1, 1, 2, 2] and e.tolist() == [0, 1, 1, 2]
    # a star: the first edge matches, every other leaf stays single (numbered
    # after the pair, in node order), repeated contracted edges are dropped
    ns, pm, e = O.contract_level(5, [0, 1, 0, 2, 0, 3, 0, 4, 2, 0])
    assert ns == 4 and pm.tolist() == [0, 0, 1, 2, 3] and e.tolist() == [0, 1, 0, 2, 0, 3]
    # orientation kept, self loops and out-of-range ids skipped
    ns, pm, e = O.contract_level(4, [1, 0, 3, 3, 9, 1, 3, 2, 2, 0])
    assert pm.tolist() == [0, 0, 1, 1] and e.tolist() == [1, 0]


def test_stop_rules():
    # a star of 600 leaves: one pair, the rest single -> ratio < 1.6 stops after
    # one contracted level
    pairs = np.stack([np.zeros(600, np.uint32), np.arange(1, 601, dtype=np.uint32)], 1).reshape(-1)
    lv = O.coarsen(601, pairs, 6, 10)
    assert [l["n"] for l in lv] == [601, 600]
    # at or below the target size: no contraction
    assert len(O.coarsen(500, pairs[:998], 6, 500)) == 1
    # max_levels caps the depth
    path = np.stack([np.arange(4095), np.arange(1, 4096)], 1).astype(np.uint32).reshape(-1)
    assert len(O.coarsen(4096, path, 3, 2)) == 3


def test_cascade_invariants_on_sbm():
    off, nb, _ = workloads.sbm_2000()
    n = len(off) - 1
    rows = np.repeat(np.arange(n), np.diff(off))
    e = np.stack([rows[rows < nb], nb[rows < nb]], 1).astype(np.uint32).reshape(-1)
    lv = O.coarsen(n, e, 6, 100)
    assert len(lv) >= 3
    for fine, coarse in zip(lv[:-1], lv[1:]):
        pm = coarse["parent_map"]
        assert len(pm) == fine["n"] and pm.max() == coarse["n"] - 1
        assert np.bincount(pm).max() <= 2  # a matching: at most two children
        ce = coarse["edges"].reshape(-1, 2)
        assert (ce[:, 0] != ce[:, 1]).all()
        key = np.minimum(ce[:, 0], ce[:, 1]).astype(np.int64) * coarse["n"] + np.maximum(ce[:, 0], ce[:, 1])
        assert len(np.unique(key)) == len(key)
        # every fine edge between different super-nodes appears once
        fe = fine["edges"].reshape(-1, 2)
        a, b = pm[fe[:, 0]], pm[fe[:, 1]]
        want = np.unique(np.minimum(a, b)[a != b].astype(np.int64) * coarse["n"] + np.maximum(a, b)[a != b])
        assert np.array_equal(np.sort(key), want)
        # maximal: no fine edge joins two single children
        single = np.bincount(pm, minlength=coarse["n"])[pm] == 1
        assert not (single[fe[:, 0]] & single[fe[:, 1]] & (fe[:, 0] != fe[:, 1])).any()


def test_fold_up_is_the_children_mean():
    lv = [{"n": 4, "edges": None, "parent_map": None},
          {"n": 2, "edges": None, "parent_map": np.array([1, 0, 1, 0], np.uint32)}]
    xyz = np.array([0, 0, 0, 1, 1, 1, 2, 4, 6, 3, 5, 7], np.float32)
    assert O.fold_up(lv, xyz).tolist() == [2.0, 3.0, 4.0, 1.0, 2.0, 3.0]


def _xorshift_star_u(seed, count):
    """Vigna's xorshift64* with exact integers: s ^= s >> 12; s ^= s << 25;
    s ^= s >> 27 (mod 2^64); output s * 2685821657736338717 mod 2^64."""
    M = (1 << 64) - 1
    s = (seed + 0x9E3779B97F4A7C15) & M
    out = []
    for _ in range(count):
        s ^= s >> 12
        s ^= (s << 25) & M
        s ^= s >> 27
        v = (s * 2685821657736338717) & M
        out.append(np.float32(np.float32(np.float64(v >> 11) / 2.0 ** 53) * np.float32(2.0) - np.float32(1.0)))
    return np.array(out, np.float32)


@pytest.mark.parametrize("seed", [0, 7, 0x5EED + 3, 0xFFFFFFFF])
def test_prolong_jitter_stream(seed):
    n = 4000
    pm = np.arange(n, dtype=np.uint32) // 2
    parent = np.zeros(3 * (n // 2), np.float32)
    got = O.prolong(parent, pm, n, 1.0, seed)
    assert np.array_equal(got, _xorshift_star_u(seed, 3 * n))
    assert got.min() >= -1.0 and got.max() < 1.0 and abs(float(got.mean())) < 0.03
    # child = parent + u * jitter in f32
    par = np.arange(3 * (n // 2), dtype=np.float32)
    j = np.float32(15.0)
    want = par.reshape(-1, 3)[pm].reshape(-1) + got * j
    assert np.array_equal(O.prolong(par, pm, n, 15.0, seed), want.astype(np.float32))


def test_sweep_schedules():
    for lvl in range(6):
        assert O.sweeps_for_level("uniform", lvl, 30) == 30
        assert O.sweeps_for_level("linear", lvl, 30) == 30 * (lvl + 1)
        assert O.sweeps_for_level("geometric", lvl, 30) == 30 * 2 ** lvl
        for s in ("uniform", "linear", "geometric"):  # coarser never fewer (multilevel.rs tests)
            assert O.sweeps_for_level(s, lvl + 1, 7) >= O.sweeps_for_level(s, lvl, 7)
    assert O.sweeps_for_level("linear", 3, 0) == 4
    assert O.sweeps_for_level("geometric", 40, 1 << 20) == 0xFFFFFFFF
