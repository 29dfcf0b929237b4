"""The analytics CLIs on the GPU.  graph-pagerank: the JSON ranks it prints equal, bit for bit,
the oracle's pull PageRank run for the same fixed iteration count (the
reference CLI has no tolerance), for stdout and --out; a graph with a
dangling node ends with the GPU error (no CPU fallback).  graph-bfs and
graph-cc print the oracle's distances and labels exactly."""
import json
import os
import subprocess

import numpy as np
import pytest

import jumpcannon as jc
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(jc.LIB_PATH), "graph-pagerank")


def _run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=300)


def test_ranks_bit_exact_with_oracle(tmp_path):
    off, nb = O.ba_graph(3000, 3, 21)
    f = tmp_path / "ba3000.bin"
    f.write_bytes(jc.csr_to_bin(off, nb))
    want, it, _ = O.pagerank_pull(off, nb, 0.85, 40, 0.0)
    assert it == 40
    r = _run(str(f), "--iters", "40")
    assert r.returncode == 0, r.stderr
    assert "backend=gpu" in r.stderr
    got = np.array(json.loads(r.stdout), dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    out = tmp_path / "ranks.json"
    r = _run(str(f), "--iters", "40", "--damping", "0.85", "--out", str(out))
    assert r.returncode == 0 and r.stdout == ""
    got2 = np.array(json.loads(out.read_text()), dtype=np.float32)
    assert np.array_equal(got2.view(np.uint32), want.view(np.uint32))


def test_dangling_node_fails_without_fallback(tmp_path):
    off, nb = O.csr_from_pairs(5, np.array([0, 1, 1, 2, 2, 3], np.uint32))  # node 4 isolated
    f = tmp_path / "dangling.bin"
    f.write_bytes(jc.csr_to_bin(off, nb))
    r = _run(str(f))
    assert r.returncode == 1
    assert "GPU PageRank failed" in r.stderr


BFS = os.path.join(os.path.dirname(jc.LIB_PATH), "graph-bfs")
CC = os.path.join(os.path.dirname(jc.LIB_PATH), "graph-cc")


def test_bfs_cli_distances_equal_oracle(tmp_path):
    off, nb = O.sbm_graph(4000, 8, 0.004, 0.00005, 11)  # several components
    f = tmp_path / "sbm.bin"
    f.write_bytes(jc.csr_to_bin(off, nb))
    want = O.bfs(off, nb, 5)
    assert (want == 0xFFFFFFFF).any()
    r = subprocess.run([BFS, str(f), "--source", "5"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "backend=gpu" in r.stderr, r.stderr
    assert json.loads(r.stdout) == want.tolist()
    out = tmp_path / "d.json"
    r = subprocess.run([BFS, str(f), "--source", "5", "--out", str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and r.stdout == "" and json.loads(out.read_text()) == want.tolist()


def test_cc_cli_labels_equal_oracle(tmp_path):
    off, nb = O.sbm_graph(4000, 8, 0.004, 0.00005, 11)
    f = tmp_path / "sbm.bin"
    f.write_bytes(jc.csr_to_bin(off, nb))
    want = O.connected_components(off, nb)
    assert len(np.unique(want)) > 1
    r = subprocess.run([CC, str(f)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "backend=gpu" in r.stderr, r.stderr
    assert json.loads(r.stdout) == want.tolist()
