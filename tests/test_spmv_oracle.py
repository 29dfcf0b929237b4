"""The SpMV oracle (cpu_spmv, spmv.rs:88-99; gpu_spmv_f16's f16 storage,
spmv.rs:68-86 and spmv_f16.wgsl; gpu_spmv_hybrid's lane/tree order,
spmv_hybrid.wgsl:44-84) pinned by the reference tests' known answers, by
scipy's sparse product on exactly representable inputs, by a hand-worked row
whose two summation orders round differently, and by numpy's float16."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O
import spmv_cases as S

VARIANTS = [O.spmv, O.spmv_hybrid, O.spmv_f16]


@pytest.mark.parametrize("fn", VARIANTS, ids=["serial", "hybrid", "f16"])
def test_identity_and_small_known_matrix(fn):
    off, nb, w = S.identity(8)
    x = np.array([1.5, -2.0, 3.0, 0.0, 7.0, -1.0, 4.0, 9.0], np.float32)
    assert np.array_equal(fn(off, nb, w, x), x)
    off, nb, w, x = S.small_known()
    assert fn(off, nb, w, x).tolist() == [5.0, 6.0, 13.0]


@pytest.mark.parametrize("fn", VARIANTS, ids=["serial", "hybrid", "f16"])
@pytest.mark.parametrize("seed", range(6))
def test_exact_integer_products_equal_scipy(fn, seed):
    """Small integers: every product and partial sum is exact in f32 (and the
    inputs exact in f16), so any order gives scipy's float64 product."""
    rng = np.random.default_rng(seed)
    n = 300
    lengths = rng.integers(0, 150, n)
    lengths[:3] = [0, 64, 129]
    off, nb, _, _ = S.ragged_rows(lengths, n, seed)
    w = rng.integers(-20, 21, len(nb)).astype(np.float32)
    x = rng.integers(-20, 21, n).astype(np.float32)
    want = sp.csr_matrix((w.astype(np.float64), nb, off), shape=(n, n)) @ x.astype(np.float64)
    assert np.array_equal(fn(off, nb, w, x).astype(np.float64), want)


def test_order_sensitive_row_hand_worked():
    off, nb, w, x = S.order_sensitive_row()
    assert O.spmv(off, nb, w, x).tolist() == [2.0 ** 24]
    assert O.spmv_hybrid(off, nb, w, x).tolist() == [2.0 ** 24 + 2.0]


def test_hybrid_equals_serial_on_rows_of_one_lane_entry():
    """Rows of at most one entry per lane whose partials are exact integers."""
    off, nb, _, x = S.ragged_rows([1, 5, 64, 0, 17], 40, 3)
    w = np.ones(len(nb), np.float32)
    x = np.round(x * 8).astype(np.float32)
    assert np.array_equal(O.spmv_hybrid(off, nb, w, x), O.spmv(off, nb, w, x))


def test_f16_known_values():
    cases = {1.0 / 3.0: 0x3555, 0.1: 0x2E66, 65504.0: 0x7BFF, 65519.0: 0x7BFF,
             65520.0: 0x7C00, -65520.0: 0xFC00, 2.0 ** -24: 0x0001, 2.0 ** -25: 0x0000,
             1.5 * 2.0 ** -25: 0x0001, 2.0 ** -14: 0x0400, -0.0: 0x8000, float("inf"): 0x7C00,
             1.0 + 2.0 ** -11: 0x3C00, 1.0 + 3 * 2.0 ** -11: 0x3C02}
    got = O.f32_to_f16_bits(np.array(list(cases), np.float32))
    assert [int(g) for g in got] == list(cases.values())


def test_f16_conversion_equals_numpy_float16():
    rng = np.random.default_rng(0)
    mags = 2.0 ** rng.uniform(-30, 17, 20000)
    vals = (mags * rng.choice([-1, 1], 20000)).astype(np.float32)
    # ties: f16 mantissa halfway points in normal and subnormal range
    ties = np.array([1.0 + (2 * k + 1) * 2.0 ** -11 for k in range(64)] +
                    [(2 * k + 1) * 2.0 ** -25 for k in range(64)], np.float32)
    vals = np.concatenate([vals, ties, -ties])
    with np.errstate(over="ignore"):  # values past 65520 go to inf, as intended
        want = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(O.f32_to_f16_bits(vals), want)
    bits = np.arange(0, 1 << 16, 7, dtype=np.uint32).astype(np.uint16)
    bits = bits[(bits & 0x7C00) != 0x7C00]  # finite values
    assert np.array_equal(O.f16_bits_to_f32(bits), bits.view(np.float16).astype(np.float32))


def test_f16_product_within_the_reference_tolerance():
    """gpu_spmv.rs f16_matches_f32_within_tolerance: within 3e-2 of f32, not equal."""
    off, nb, w, x = S.f16_fractional()
    dev = np.abs(O.spmv_f16(off, nb, w, x) - O.spmv(off, nb, w, x))
    assert dev.max() < 3e-2 and dev.max() > 0.0


@pytest.mark.parametrize("seed", range(8))
def test_random_matrices_within_the_reference_tolerance(seed):
    """gpu_equals_cpu_spmv / hybrid_equals_baseline_and_cpu: 1e-4 against f64."""
    off, nb, w, x = S.random_sparse(1 + seed * 5, seed * 977)
    n = len(off) - 1
    want = sp.csr_matrix((w.astype(np.float64), nb, off), shape=(n, n)) @ x.astype(np.float64)
    assert np.abs(O.spmv(off, nb, w, x) - want).max(initial=0.0) < 1e-4
    assert np.abs(O.spmv_hybrid(off, nb, w, x) - want).max(initial=0.0) < 1e-4


def test_power_law_hubs_within_tolerance():
    off, nb, w, x = S.power_law(200, 3, 42)
    assert off[1] - off[0] == 200
    want = sp.csr_matrix((w.astype(np.float64), nb, off), shape=(200, 200)) @ x.astype(np.float64)
    assert np.abs(O.spmv_hybrid(off, nb, w, x) - want).max() < 1e-4
    assert np.abs(O.spmv(off, nb, w, x) - want).max() < 1e-4
