"""CPU checks of the C ABI (no GPU compute): the library loads and exports every
entry point include/jc.h declares, the no-device path errs like the reference,
settings decode with serde semantics, and the product's fixture generators
agree with the oracle's restatement."""
import ctypes
import os
import re

import numpy as np
import pytest

import jumpcannon as jc
from jumpcannon import workloads
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "jc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(jc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(jc.LIB_PATH)
    syms = _declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the Python front-end binds all of them
    assert set(syms) <= set(jc.SIGNATURES)


def test_no_device_init_errs_like_reference():
    """fa2_bh.rs:228-231: init without a GPU returns an error (caller falls back)."""
    if jc.device_count() > 0:
        pytest.skip("a device is visible")
    e = jc.Fa2BhEngine()
    with pytest.raises(jc.NoDeviceError):
        e.init(np.array([0, 1, 2], np.uint32), np.array([1, 0], np.uint32), np.zeros(6, np.float32))
    with pytest.raises(jc.JcError) as ei:
        e.step(1)
    assert ei.value.code == jc.JC_E_STATE


def test_descriptor():
    d = jc.Fa2BhEngine().descriptor
    assert d["id"] == "fa2-bh-hip" and d["kind"] == "physics"
    assert d["needs_edges"] and d["needs_cpu_positions"] and not d["needs_gpu_positions_buffer"]


def test_set_params_serde_semantics():
    """fa2_bh.rs:212-220: null keeps; typed decode errors; unknown keys ignored."""
    e = jc.Fa2BhEngine()
    e.set_params(None)
    e.set_params("null")
    e.set_params({"theta": 1.2, "gravity": 3, "strong_gravity": True, "extra": {"x": [1, 2]}})
    for bad in ({"theta": "1.0"}, {"strong_gravity": 1}, {"gravity": None}, [1, 2], "{",
                '{"theta": 1.0,}', "3"):
        with pytest.raises(jc.JcError) as ei:
            e.set_params(bad)
        assert ei.value.code == jc.JC_E_INVALID


def test_default_settings_match_reference_defaults():
    """Fa2BhSettings::default (fa2_bh.rs:109-125)."""
    s = jc.default_fa2bh_settings()
    assert (s.gravity, s.strong_gravity, s.scaling_ratio, s.edge_weight_influence,
            s.jitter_tolerance, s.lin_log_mode, s.prevent_overlap, s.time_step,
            round(s.theta, 6), s.max_displacement) == (1.0, 0, 2.0, 1.0, 1.0, 0, 0, 1.0, 0.7, 10.0)
    g = jc.SgdSettings()
    jc.lib().jc_sgd_default_settings(ctypes.addressof(g))
    assert (g.n_pivots, g.sweeps_per_step, g.n_anneal_steps, g.seed) == (50, 1, 30, 0x5EEDC0DE12345678)
    assert abs(g.eta_max - 1.0) < 1e-7 and abs(g.eta_min - 0.01) < 1e-7


def test_brute_engine_descriptor_settings_and_params():
    """fa2_brute.rs:63-93 (Fa2Settings and its defaults), :143-158 (descriptor),
    :170-178 (set_params: null keeps, decode errors are errors, unknown keys
    such as fa2-bh's theta are ignored)."""
    s = jc.default_fa2_settings()
    assert (s.gravity, s.strong_gravity, s.scaling_ratio, s.edge_weight_influence,
            s.jitter_tolerance, s.lin_log_mode, s.prevent_overlap, s.time_step,
            s.max_displacement) == (1.0, 0, 2.0, 1.0, 1.0, 0, 0, 1.0, 10.0)
    e = jc.Fa2BruteEngine()
    d = e.descriptor
    assert d["id"] == "fa2-brute-hip" and d["kind"] == "physics"
    assert d["needs_edges"] and d["needs_cpu_positions"] and not d["needs_gpu_positions_buffer"]
    e.set_params(None)
    e.set_params({"theta": "not a number here", "gravity": 3, "lin_log_mode": True})
    for bad in ({"gravity": "1.0"}, {"strong_gravity": 1}, {"max_displacement": None}, "[1]"):
        with pytest.raises(jc.JcError) as ei:
            e.set_params(bad)
        assert ei.value.code == jc.JC_E_INVALID
    if jc.device_count() == 0:
        with pytest.raises(jc.NoDeviceError):
            e.init(np.array([0, 1, 2], np.uint32), np.array([1, 0], np.uint32), np.zeros(6, np.float32))


@pytest.mark.parametrize("n,m,seed", [(2, 1, 0), (50, 3, 1), (5000, 4, 0xC0FFEE)])
def test_ba_generator_matches_oracle(n, m, seed):
    a = jc.ba_graph(n, m, seed)
    b = O.ba_graph(n, m, seed)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_sbm_rgg_cube_generators_match_oracle():
    assert np.array_equal(jc.sbm_pairs(2000, 4, 0.02, 0.001, 5), O.sbm_pairs(2000, 4, 0.02, 0.001, 5))
    xa, pa = jc.rgg_pairs(3000, 0.03, 9)
    xb, pb = O.rgg_pairs(3000, 0.03, 9)
    assert np.array_equal(xa, xb) and np.array_equal(pa, pb)
    assert np.array_equal(jc.uniform_cube(777, 10.0, 3), O.uniform_cube(777, 10.0, 3))


def test_sbm_2000_workload_shape():
    off, nb, xyz = workloads.sbm_2000()
    n = len(off) - 1
    assert n == 2000 and xyz.shape == (6000,) and np.abs(xyz).max() <= 1.0
    # E|E| ~ 4*C(500,2)*0.02 + 6*500*500*0.001 = 9980 + 1500
    assert 10500 < len(nb) // 2 < 12500


def _path_csr(n):
    """CsrGraph::path(n) (sim.rs:33-51): 0-1-2-...-(n-1)."""
    off, nb = [0], []
    for i in range(n):
        if i > 0:
            nb.append(i - 1)
        if i + 1 < n:
            nb.append(i + 1)
        off.append(len(nb))
    return np.array(off, np.uint32), np.array(nb, np.uint32)


def test_csr_bin_roundtrip(tmp_path):
    """sim.rs:135-149: write_bin then load_bin gives the same graph; the bytes are
    the little-endian header plus both arrays."""
    off, nb = _path_csr(8)
    p = tmp_path / "path8.bin"
    jc.write_csr_bin(p, off, nb)
    raw = p.read_bytes()
    want = np.concatenate([np.array([8, len(nb)], "<u4"), off.astype("<u4"), nb.astype("<u4")])
    assert raw == want.tobytes()
    o2, n2 = jc.load_csr_bin(p)
    assert np.array_equal(o2, off) and np.array_equal(n2, nb)
    bo, bn = jc.ba_graph(3000, 3, 11)
    o3, n3 = jc.csr_from_bin(jc.csr_to_bin(bo, bn))
    assert np.array_equal(o3, bo) and np.array_equal(n3, bn)
    o4, n4 = jc.csr_from_bin(jc.csr_to_bin(np.zeros(1, np.uint32), np.zeros(0, np.uint32)))
    assert list(o4) == [0] and len(n4) == 0


def test_csr_bin_rejects_truncated_and_length_mismatch():
    """sim.rs:151-179: a 4-byte file is truncated; a header claiming 4 nodes and 8
    edges with no payload is a length mismatch."""
    with pytest.raises(jc.JcError) as ei:
        jc.csr_from_bin(bytes(4))
    assert ei.value.code == jc.JC_E_INVALID and "truncated" in str(ei.value)
    hdr = np.array([4, 8], "<u4").tobytes()
    with pytest.raises(jc.JcError) as ei:
        jc.csr_from_bin(hdr)
    assert "length mismatch" in str(ei.value) and "expected 60" in str(ei.value)
    with pytest.raises(jc.JcError):
        jc.csr_from_bin(jc.csr_to_bin(*_path_csr(5)) + b"\0")


def test_multilevel_settings_decode_without_a_device():
    """multilevel.rs:307-331 (and its unit tests rejects_self_wrap,
    rejects_unknown_inner, accepts_known_inner): settings decode and the inner
    id checks run on the host, before any device work."""
    e = jc.MultilevelEngine()
    assert e.descriptor["id"] == "multilevel-hip"
    e.set_params({"inner": "sgd-stress-hip"})
    e.set_params({"inner": "fa2-brute-hip"})
    e.set_params({"inner": "fa2-bh-hip", "inner_params": {"theta": 0.9, "nested": [1, {"a": "}"}]},
                  "max_levels": 3, "target_size": 100, "sweeps_per_level": 7, "sweep_schedule": "geometric",
                  "spring_len": 2.5, "seed": 12345})
    e.set_params(None)
    e.set_params({"inner_params": None})  # null forwards the inner defaults
    for bad, msg in (({"inner": "multilevel"}, "cannot wrap itself"),
                     ({"inner": "fa2-bh"}, "unknown inner engine"),
                     ({"inner": 3}, "invalid type for field `inner`"),
                     ({"sweep_schedule": "Linear"}, "unknown variant"),
                     ({"seed": -1}, "invalid type"),
                     ({"max_levels": 2.5}, "invalid type"),
                     ('{"inner": "fa2-bh-hip"', "expected")):
        with pytest.raises(jc.JcError) as ei:
            e.set_params(bad)
        assert ei.value.code == jc.JC_E_INVALID and msg in str(ei.value), (bad, str(ei.value))
    with pytest.raises(jc.JcError) as ei:
        e.state()  # before init
    assert ei.value.code == jc.JC_E_STATE


def test_buffer_lengths_checked_before_the_call():
    """Every buffer the C side reads a fixed count from is checked in Python
    first (a short one would be read past its end)."""
    e = jc.Fa2BhEngine()  # not initialised: n = 0
    off2 = np.array([0, 1, 2], np.uint32)
    nb2 = np.array([1, 0], np.uint32)
    cases = [
        lambda: e.reinit(off2, nb2, np.zeros(5, np.float32)),
        lambda: e.set_state(positions=np.zeros(3, np.float32)),
        lambda: e.set_state(old_force=np.zeros(3, np.float32)),
        lambda: e.set_state(speed_eff2=np.zeros(1, np.float32)),
        lambda: e.forces_at(np.zeros(3, np.float32)),
        lambda: jc.StressJudge(10, np.array([0], np.uint32), np.zeros(9, np.uint16)),
    ]
    for k, call in enumerate(cases):
        with pytest.raises(jc.JcError) as ei:
            call()
        assert ei.value.code == jc.JC_E_INVALID and "length" in str(ei.value), k
    # the network's buffers, checked without creating a device handle
    m = jc.Mpnn.__new__(jc.Mpnn)
    m._h, m.n, m.P = None, 4, int(jc.lib().jc_mpnn_param_count())
    for call in (lambda: m.set_params(np.zeros(m.P - 1, np.float32)),
                 lambda: m.set_target(np.zeros(11, np.float32)),
                 lambda: m.set_graph(np.array([0, 1, 2, 3, 4], np.uint32), np.array([1, 0, 3, 2], np.uint32),
                                     noise=np.zeros(11, np.float32))):
        with pytest.raises(jc.JcError) as ei:
            call()
        assert ei.value.code == jc.JC_E_INVALID


def test_batch_schedule_arguments_checked_without_a_device():
    """The batch schedule API validates its arguments on the host: unknown modes
    and workgroup sizes are JC_E_INVALID, the persistent schedule is switched
    off, and the phased / autotuned ones are accepted before init."""
    b = jc.Fa2BhBatch(theta=1.0)
    b.set_schedule("phased")
    b.set_schedule("auto")
    with pytest.raises(jc.JcError):
        b.set_schedule("persistent", 256)
    with pytest.raises(jc.JcError):
        b.set_schedule("nonsense")
    rc = jc.lib().jc_fa2bh_batch_set_schedule(b._h, 1, 300)
    assert rc == jc.JC_E_INVALID
    assert b.schedule()[0] == "auto"


def test_batch_tuning_knobs_checked_without_a_device():
    """Knobs of the phased batch schedule are validated on the host."""
    b = jc.Fa2BhBatch(theta=1.0)
    b.set_tuning(force_block=512, com_depth=4, com_chunk=2048, sort_key_bits=48)
    b.set_tuning()
    for bad in (dict(force_block=100), dict(com_depth=2), dict(com_depth=6), dict(com_chunk=512),
                dict(sort_key_bits=40), dict(sort_key_bits=80)):
        with pytest.raises(jc.JcError):
            b.set_tuning(**bad)
    assert b.schedule()[0] == "auto" and len(b.schedule()) == 6
