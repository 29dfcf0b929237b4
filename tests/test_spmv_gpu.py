"""SpMV on the GPU through the C ABI, bit-exact against the oracle for the
three reference variants (gpu_spmv, gpu_spmv_f16, gpu_spmv_hybrid in
crates/graph-compute/src/analytics/spmv.rs): the reference tests' shapes, their
random and power-law generators, ragged row lengths around the kernels' row
limits, and the 1M-node BA graph with PageRank's per-entry weights."""
import numpy as np
import pytest

import jumpcannon as jc
from jumpcannon import workloads
from oracle import oracle as O
import spmv_cases as S

pytestmark = pytest.mark.gpu

PAIRS = [(jc.spmv, O.spmv), (jc.spmv_f16, O.spmv_f16), (jc.spmv_hybrid, O.spmv_hybrid)]
IDS = ["f32", "f16", "hybrid"]


def _eq(fn, ref, off, nb, w, x):
    got = fn(off, nb, w, x)
    exp = ref(off, nb, w, x)
    assert got.dtype == np.float32 and got.shape == exp.shape
    bad = np.nonzero(got.view(np.uint32) != exp.view(np.uint32))[0]
    assert bad.size == 0, (bad[:8], got[bad[:8]], exp[bad[:8]])
    return got


@pytest.mark.parametrize("fn,ref", PAIRS, ids=IDS)
def test_reference_shapes(fn, ref):
    off, nb, w = S.identity(8)
    x = np.array([1.5, -2.0, 3.0, 0.0, 7.0, -1.0, 4.0, 9.0], np.float32)
    assert np.array_equal(_eq(fn, ref, off, nb, w, x), x)
    assert _eq(fn, ref, *S.small_known()).tolist() == [5.0, 6.0, 13.0]
    _eq(fn, ref, *S.f16_fractional())


@pytest.mark.parametrize("fn,ref", PAIRS, ids=IDS)
@pytest.mark.parametrize("seed", range(24))
def test_random_sparse(fn, ref, seed):
    _eq(fn, ref, *S.random_sparse(1 + (seed * 7) % 39, seed * 409 + 5))


@pytest.mark.parametrize("fn,ref", PAIRS, ids=IDS)
def test_power_law_hubs_and_ragged_rows(fn, ref):
    off, nb, w, x = S.power_law(200, 3, 42)
    _eq(fn, ref, off, nb, w, x)
    _eq(fn, ref, *S.power_law(3000, 2, 7))
    # around the light-row limit (32), the wave width, the long-row round (1024)
    lengths = [0, 1, 31, 32, 33, 63, 64, 65, 127, 128, 129, 1023, 1024, 1025, 2049, 5000] * 5
    _eq(fn, ref, *S.ragged_rows(lengths, 7000, 11))


def test_order_sensitive_row():
    off, nb, w, x = S.order_sensitive_row()
    assert jc.spmv(off, nb, w, x).tolist() == [2.0 ** 24]
    assert jc.spmv_f16(off, nb, np.array([1.0] * 65, np.float32), x).tolist() == [65.0]
    assert jc.spmv_hybrid(off, nb, w, x).tolist() == [2.0 ** 24 + 2.0]


def test_f16_special_values():
    """Overflow to inf, subnormal f16 weights, negative zero products."""
    w = np.array([70000.0, 2.0 ** -24, 3 * 2.0 ** -26, -0.0, 1e-8], np.float32)
    off = np.array([0, 1, 2, 3, 4, 5], np.uint32)
    nb = np.arange(5, dtype=np.uint32)
    x = np.array([1.0, 1.0, 2.0, 5.0, 1.0], np.float32)
    _eq(jc.spmv_f16, O.spmv_f16, off, nb, w, x)


def test_errors_and_empty():
    off, nb, w, x = S.small_known()
    with pytest.raises(jc.JcError) as e:
        jc.spmv(off, np.array([0, 2, 1, 0, 3], np.uint32), w, x)
    assert e.value.code == jc.JC_E_INVALID
    with pytest.raises(jc.JcError):
        jc.spmv_hybrid(off, nb, w[:4], x)
    with pytest.raises(jc.JcError):
        jc.spmv_f16(off, nb, w, x[:2])
    for fn in (jc.spmv, jc.spmv_f16, jc.spmv_hybrid):
        assert fn(np.zeros(1, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.float32),
                  np.zeros(0, np.float32)).size == 0


def test_plan_reuse_and_pagerank_step_as_spmv():
    """The 1M-node BA graph with w_e = 1/deg(u) (PageRank's pull step as an
    SpMV): bit-exact for all three kinds, two products on one plan."""
    off, nb, _ = workloads.ba_1m()
    n = len(off) - 1
    deg = np.diff(off).astype(np.float32)
    w = (np.float32(1.0) / deg)[nb].astype(np.float32)
    rng = np.random.default_rng(5)
    for kind, ref in ((jc.SPMV_F32, O.spmv), (jc.SPMV_F16, O.spmv_f16),
                      (jc.SPMV_HYBRID, O.spmv_hybrid)):
        plan = jc.SpmvPlan(off, nb, w, kind)
        for _ in range(2):
            x = rng.uniform(0, 2.0 / n, n).astype(np.float32)
            y, ms = plan.run(x)
            assert np.array_equal(y.view(np.uint32), ref(off, nb, w, x).view(np.uint32))
        print(f"kind {kind}: product {ms:.3f} ms, {plan.time_products(20):.4f} ms resident")
    # one pull step: teleport + d * (A x) with x = 1/n equals the oracle's first iterate
    x = np.full(n, np.float32(1.0) / np.float32(n), np.float32)
    y = jc.spmv(off, nb, w, x)
    d = np.float32(0.85)
    step = (np.float32(1.0) - d) * (np.float32(1.0) / np.float32(n)) + d * y
    r1, _, _ = O.pagerank_pull(off, nb, 0.85, max_iters=1, tol=0.0)
    assert np.abs(step - r1).max() <= 1e-6 * np.abs(r1).max()
