"""PageRank on the GPU against the oracle's pull restatement (pagerank.wgsl:30-46):
ranks bit-exact (every row is summed in CSR order, hubs included), the exact
L1 residual and the iteration count identical; plus the reference's own
closed-form and scale checks."""
import numpy as np
import pytest

import jumpcannon as jc
from jumpcannon import workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _ring(n, k=1):
    pairs = [[i, (i + j) % n] for i in range(n) for j in range(1, k + 1)]
    return O.csr_from_pairs(n, np.array(pairs, np.uint32).reshape(-1))


def _path(n):
    return O.csr_from_pairs(n, np.array([[i, i + 1] for i in range(n - 1)], np.uint32).reshape(-1))


def _check_exact(off, nb, iters, tol):
    r, it, res = jc.pagerank(off, nb, 0.85, iters, tol)
    ro, ito, reso = O.pagerank_pull(off, nb, 0.85, iters, tol)
    assert it == ito
    assert res == reso
    assert np.array_equal(r.view(np.uint32), ro.view(np.uint32)), np.abs(r - ro).max()
    return r, it, res


@pytest.mark.parametrize("name", ["path16", "ring32"])
def test_matches_cpu_oracle(name):
    """pagerank.rs:355-382: GPU == CPU within 1e-4 and mass ~ 1 (here: bit-exact pull)."""
    off, nb = _path(16) if name == "path16" else _ring(32)
    r, _, _ = _check_exact(off, nb, 50, 0.0)
    push = O.pagerank_push(off, nb, 0.85, 50)
    assert np.abs(r - push).max() < 1e-4
    assert abs(float(r.sum()) - 1.0) < 1e-2


def test_star_closed_form():
    L, d = 40, 0.85
    off, nb = O.csr_from_pairs(L + 1, np.array([[0, i] for i in range(1, L + 1)], np.uint32).reshape(-1))
    r, _, _ = _check_exact(off, nb, 300, 1e-12)
    n = L + 1
    rc = (1 + d * L) / (n * (1 + d))
    rl = (1 - d) / n + d / L * rc
    assert abs(r[0] - rc) < 1e-3 and np.abs(r[1:] - rl).max() < 1e-3


def test_hub_rows_bit_exact():
    off, nb = O.ba_graph(30000, 5, 17)
    assert np.diff(off).max() > 96  # exercises the hub kernel
    _check_exact(off, nb, 40, 0.0)


def test_rejects_dangling():
    """pagerank.rs:383-397."""
    with pytest.raises(jc.JcError) as ei:
        jc.pagerank(np.array([0, 1, 1], np.uint32), np.array([1], np.uint32))
    assert ei.value.code == jc.JC_E_DANGLING


def test_ba_1m_converges_bit_exact():
    """The 1M-node BA CSR, d=0.85, L1 tol 1e-6, at most 100 iterations."""
    off, nb, _ = workloads.ba_1m()
    r, it, res = _check_exact(off, nb, 100, 1e-6)
    assert res < 1e-6 or it == 100
    assert abs(float(r.astype(np.float64).sum()) - 1.0) < 1e-3


def test_kring_2m_uniform():
    """gpu_pagerank_scale.rs:40-92: 2M-node k-ring (k=4) is uniform within 1% of 1/n."""
    n, k = 2_000_000, 4
    i = np.arange(n, dtype=np.int64)
    nb = np.stack([(i + j) % n for j in (-4, -3, -2, -1, 1, 2, 3, 4)], axis=1)
    nb = np.sort(nb, axis=1).astype(np.uint32).reshape(-1)
    off = (np.arange(n + 1, dtype=np.int64) * 8).astype(np.uint32)
    plan = jc.PageRankPlan(off, nb)
    r, it, res, ms = plan.run(30, 0.0)
    assert np.abs(r - 1.0 / n).max() < 0.01 / n
    assert abs(float(r.astype(np.float64).sum()) - 1.0) < 1e-3


def test_plan_reuse_is_deterministic():
    off, nb = O.ba_graph(50000, 4, 3)
    plan = jc.PageRankPlan(off, nb)
    a = plan.run(100, 1e-7)
    b = plan.run(100, 1e-7)
    assert a[1] == b[1] and a[2] == b[2]
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    assert plan.time_iterations(20) > 0
