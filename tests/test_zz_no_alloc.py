"""jc_pagerank_noalloc (PageRank over a device-resident CSR with its workspace in
a static device array) against jc_pagerank on the same inputs: ranks bit for
bit, iteration count and residual equal, on consecutive calls whose second
input is smaller, so whatever the first call leaves in the shared workspace
must not leak into the second."""
import numpy as np
import pytest

import jumpcannon as jc
from jumpcannon import workloads
from oracle import oracle as O
from hip_buffers import DeviceArray

pytestmark = pytest.mark.gpu


def _noalloc(off, nb, **kw):
    off = np.ascontiguousarray(off, np.uint32)
    nb = np.ascontiguousarray(nb, np.uint32)
    n = len(off) - 1
    d_off, d_nb = DeviceArray(off), DeviceArray(nb if len(nb) else np.zeros(1, np.uint32))
    d_out = DeviceArray(dtype=np.float32, count=max(n, 1))
    it, res = jc.pagerank_noalloc(n, d_off.ptr, d_nb.ptr, len(nb), d_out.ptr, **kw)
    return d_out.numpy()[:n], it, res


def _same(off, nb, **kw):
    got, it, res = _noalloc(off, nb, **kw)
    want, it2, res2 = jc.pagerank(off, nb, **kw)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), np.abs(got - want).max()
    assert it == it2 and res == res2
    return it


def test_noalloc_equals_pagerank_on_shrinking_inputs():
    off, nb, _ = workloads.ba_1m()
    it = _same(off, nb)
    print(f"BA 1M: {it} iterations")
    off2, nb2 = O.ba_graph(200_000, 4, 17)     # smaller, hubs above the 32-entry row limit
    _same(off2, nb2)
    _same(off2, nb2, tol_l1=-1.0, max_iters=37)  # odd count: the ranks end in the second buffer
    off3, nb3 = O.ba_graph(3000, 3, 8)
    _same(off3, nb3)
    _same(off3, nb3, damping=0.5, tol_l1=1e-9)


def test_noalloc_errors_and_empty():
    with pytest.raises(jc.JcError) as e:
        jc.pagerank_noalloc(jc.PAGERANK_NOALLOC_MAX_N + 1, 0, 0, 0, 0)
    assert e.value.code == jc.JC_E_INVALID and "at most" in str(e.value)
    assert jc.pagerank_noalloc(0, 0, 0, 0, 0) == (0, 0.0)
    off = np.array([0, 1, 1, 2], np.uint32)  # node 1 has no neighbour
    nb = np.array([2, 0], np.uint32)
    with pytest.raises(jc.JcError) as e:
        _noalloc(off, nb)
    assert e.value.code == jc.JC_E_DANGLING and "node 1" in str(e.value)
    with pytest.raises(jc.JcError) as e:
        _noalloc(np.array([0, 1, 2], np.uint32), np.array([1, 5], np.uint32))
    assert e.value.code == jc.JC_E_INVALID and "out of range" in str(e.value)
    with pytest.raises(jc.JcError) as e:
        _noalloc(np.array([0, 2, 1], np.uint32), np.array([1, 0], np.uint32))
    assert e.value.code == jc.JC_E_INVALID
    # a good call after the failed ones
    off3, nb3 = O.ba_graph(500, 2, 3)
    _same(off3, nb3)
