"""graph-pagerank CLI (the reference's crates/graph-compute/src/bin/pagerank.rs
over the C ABI): argument handling and error paths that need no GPU.  The
bit-exact GPU run is in test_cli_gpu.py."""
import os
import subprocess

import numpy as np
import pytest

import jumpcannon as jc
from oracle import oracle as O

CLI = os.path.join(os.path.dirname(jc.LIB_PATH), "graph-pagerank")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="graph-pagerank not built")


def _run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)


def test_help_and_usage_errors():
    r = _run("--help")
    assert r.returncode == 0 and "graph-pagerank <csr.bin>" in r.stderr
    r = _run()
    assert r.returncode == 2 and "missing required <csr.bin> path" in r.stderr
    r = _run("g.bin", "--iters", "-3")
    assert r.returncode == 2 and "--iters" in r.stderr
    r = _run("g.bin", "--damping")
    assert r.returncode == 2 and "--damping needs a value" in r.stderr
    r = _run("g.bin", "extra.bin")
    assert r.returncode == 2 and "unexpected argument" in r.stderr


def test_cpu_flag_is_rejected():
    """No CPU path in this build: --cpu fails loudly instead of running a fallback."""
    r = _run("g.bin", "--cpu")
    assert r.returncode == 2 and "GPU only" in r.stderr


def test_missing_and_truncated_files(tmp_path):
    r = _run(str(tmp_path / "absent.bin"))
    assert r.returncode == 1 and "loading CSR" in r.stderr
    bad = tmp_path / "short.bin"
    bad.write_bytes(b"\x03\x00\x00")
    r = _run(str(bad))
    assert r.returncode == 1 and "CSR truncated" in r.stderr


@pytest.mark.skipif(jc.device_count() > 0, reason="checks the no-device error")
def test_no_device_is_an_error(tmp_path):
    off, nb = O.csr_from_pairs(4, np.array([0, 1, 1, 2, 2, 3, 3, 0], np.uint32))
    f = tmp_path / "ring4.bin"
    f.write_bytes(jc.csr_to_bin(off, nb))
    r = _run(str(f))
    assert r.returncode == 1 and "GPU PageRank failed" in r.stderr


# ---- graph-bfs / graph-cc (crates/graph-compute/src/bin/bfs.rs, cc.rs) ----
BFS = os.path.join(os.path.dirname(jc.LIB_PATH), "graph-bfs")
CC = os.path.join(os.path.dirname(jc.LIB_PATH), "graph-cc")


def _run_bin(path, *args):
    return subprocess.run([path, *args], capture_output=True, text=True, timeout=120)


def _ring_file(tmp_path):
    off, nb = O.csr_from_pairs(4, np.array([0, 1, 1, 2, 2, 3, 3, 0], np.uint32))
    f = tmp_path / "ring4.bin"
    f.write_bytes(jc.csr_to_bin(off, nb))
    return f


def test_bfs_cli_usage_errors(tmp_path):
    r = _run_bin(BFS, "--help")
    assert r.returncode == 0 and "graph-bfs <csr.bin> --source <u32>" in r.stderr
    r = _run_bin(BFS)
    assert r.returncode == 2 and "missing required <csr.bin> path" in r.stderr
    r = _run_bin(BFS, "g.bin")
    assert r.returncode == 2 and "missing required --source <u32>" in r.stderr
    r = _run_bin(BFS, "g.bin", "--source")
    assert r.returncode == 2 and "--source needs a value" in r.stderr
    r = _run_bin(BFS, "g.bin", "--source", "x1")
    assert r.returncode == 2 and "parsing --source" in r.stderr
    r = _run_bin(BFS, "g.bin", "--source", "0", "--cpu")
    assert r.returncode == 2 and "GPU only" in r.stderr
    r = _run_bin(BFS, str(tmp_path / "absent.bin"), "--source", "0")
    assert r.returncode == 1 and "loading CSR" in r.stderr
    # bin/bfs.rs:80-85: the range check comes before any GPU work
    r = _run_bin(BFS, str(_ring_file(tmp_path)), "--source", "4")
    assert r.returncode == 1 and "--source 4 out of range (graph has 4 nodes)" in r.stderr


def test_cc_cli_usage_errors(tmp_path):
    r = _run_bin(CC, "-h")
    assert r.returncode == 0 and "graph-cc <csr.bin> [--out labels.json]" in r.stderr
    r = _run_bin(CC)
    assert r.returncode == 2 and "missing required <csr.bin> path" in r.stderr
    r = _run_bin(CC, "g.bin", "--out")
    assert r.returncode == 2 and "--out needs a path" in r.stderr
    r = _run_bin(CC, "g.bin", "--cpu")
    assert r.returncode == 2 and "GPU only" in r.stderr
    r = _run_bin(CC, "a.bin", "b.bin")
    assert r.returncode == 2 and "unexpected argument" in r.stderr


@pytest.mark.skipif(jc.device_count() > 0, reason="checks the no-device error")
def test_bfs_cc_no_device_is_an_error(tmp_path):
    f = _ring_file(tmp_path)
    r = _run_bin(BFS, str(f), "--source", "0")
    assert r.returncode == 1 and "GPU BFS failed" in r.stderr
    r = _run_bin(CC, str(f))
    assert r.returncode == 1 and "GPU connected components failed" in r.stderr
