"""The graph_compute.Compute worker (jumpcannon.worker) on the CPU: the wire
contract, LoadGraph, Health, ListEngines and every Subscribe rejection, which
all answer before an engine is stepped.  The reference's own tests are
crates/graph-compute/tests/grpc_stream.rs and list_engines_grpc.rs; the
frame-streaming cases need the GPU (test_worker_gpu.py)."""
import grpc
import numpy as np
import pytest

import jumpcannon as jc
from jumpcannon import compute_proto as P
from jumpcannon import worker as W


def _pos_bytes(p):
    return np.asarray(p, "<f4").tobytes()


def _cycle4():
    return np.array([0, 2, 4, 6, 8], np.uint32), np.array([1, 3, 0, 2, 1, 3, 0, 2], np.uint32)


@pytest.fixture
def served():
    state = W.SimState(*W.path_graph(3))
    server, port, stop = W.serve(state, "127.0.0.1:0", start_loop=False, max_workers=4)
    ch = grpc.insecure_channel(f"127.0.0.1:{port}")
    yield state, W.ComputeClient(ch)
    ch.close()
    stop.set()
    server.stop(0.5)


def test_wire_encoding_matches_compute_proto():
    """Field numbers and types of compute.proto: bytes hand-encoded from the
    proto3 wire rules decode into the expected fields, and back."""
    # SubscribeRequest{graph_id=1:"g", layout_id=2:"fa2-bh", graph_revision=5:7, selection_generation=6:9}
    raw = bytes([0x0A, 1]) + b"g" + bytes([0x12, 6]) + b"fa2-bh" + bytes([0x28, 7, 0x30, 9])
    r = P.SubscribeRequest.FromString(raw)
    assert (r.graph_id, r.layout_id, r.graph_revision, r.selection_generation) == ("g", "fa2-bh", 7, 9)
    assert r.SerializeToString() == raw
    # PositionDelta{frame=1:3, positions=2:<4 bytes>, n_nodes=3:1, graph_revision=4:2}
    d = P.PositionDelta(frame=3, positions=b"\x00\x00\x80\x3f", n_nodes=1, graph_revision=2)
    assert d.SerializeToString() == bytes([0x08, 3, 0x12, 4, 0, 0, 0x80, 0x3F, 0x18, 1, 0x20, 2])
    # LoadGraphResponse{ok=1:true, n_nodes=2:4, error=3:"", graph_revision=4:42}
    assert P.LoadGraphResponse(ok=True, n_nodes=4, graph_revision=42).SerializeToString() == \
        bytes([0x08, 1, 0x10, 4, 0x20, 42])
    # HealthResponse fields 1..8 and the six RPCs of the service
    h = P.HealthResponse(ok=True, n_nodes=2, frame=3, requested_layout_id="a", effective_layout_id="b",
                         fallback_reason="c", graph_revision=4, selection_generation=5)
    assert P.HealthResponse.FromString(h.SerializeToString()) == h
    assert [m.name for m in P.FILE.services_by_name["Compute"].methods] == \
        ["Subscribe", "Health", "ListEngines", "TopoFisheye", "ExchangeHalo", "LoadGraph"]


def test_struct_params_turn_whole_numbers_into_integers():
    """service.rs number_value_to_json: whole doubles become integers, others stay floats."""
    s = P.Struct()
    s.fields["theta"].number_value = 0.8
    s.fields["n_pivots"].number_value = 50.0
    s.fields["lin_log_mode"].bool_value = True
    s.fields["nested"].struct_value.fields["k"].number_value = 3.0
    got = W.struct_to_params(s)
    assert got == {"theta": 0.8, "n_pivots": 50, "lin_log_mode": True, "nested": {"k": 3}}
    assert isinstance(got["n_pivots"], int) and isinstance(got["theta"], float)


def test_ring_seed_is_the_reference_lcg():
    """sim.rs ring_seed: s <- s * 1664525 + 1013904223 (mod 2^32) from 0x9E3779B1,
    each draw (f32(s) / f32(2^32 - 1)) * 2 - 1 times sqrt(n) * 5, in f32."""
    n = 5
    got = W.ring_seed(n)
    s, want = 0x9E3779B1, []
    radius = np.float32(np.sqrt(np.float32(n))) * np.float32(5.0)
    for _ in range(3 * n):
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        u = np.float32(s) / np.float32(4294967295.0)
        want.append((u * np.float32(2.0) - np.float32(1.0)) * radius)
    assert np.array_equal(got, np.array(want, np.float32))
    assert np.abs(W.ring_seed(1000)).max() <= np.float32(np.sqrt(1000.0) * 5.0)
    assert len(W.ring_seed(0)) == 0


def test_load_graph_replaces_graph_and_positions(served):
    """grpc_stream.rs load_graph_preserves_supplied_positions_and_rejects_bad_arity_atomically."""
    state, cl = served
    supplied = [-1.0, -2.0, -3.0, 4.0, 5.0, 6.0, 7.5, 8.5, 9.5]
    r = cl.LoadGraph(P.LoadGraphRequest(csr=jc.csr_to_bin(*W.path_graph(3)), positions=_pos_bytes(supplied),
                                        graph_revision=77))
    assert r.ok and r.n_nodes == 3 and r.graph_revision == 77, r.error
    assert np.array_equal(state.positions, np.array(supplied, np.float32))
    before = (state.offsets.copy(), state.neighbors.copy())
    bad = cl.LoadGraph(P.LoadGraphRequest(csr=jc.csr_to_bin(*_cycle4()), positions=_pos_bytes([0.0, 1.0, 2.0]),
                                          graph_revision=78))
    assert not bad.ok and "positions length" in bad.error and bad.graph_revision == 77
    assert state.graph_revision == 77
    assert np.array_equal(state.offsets, before[0]) and np.array_equal(state.neighbors, before[1])
    assert np.array_equal(state.positions, np.array(supplied, np.float32))


def test_load_graph_soft_errors(served):
    state, cl = served
    r = cl.LoadGraph(P.LoadGraphRequest(csr=b"\x01\x02", graph_revision=5))
    assert not r.ok and r.error.startswith("decode CSR") and r.graph_revision == 0
    csr = jc.csr_to_bin(*_cycle4())
    r = cl.LoadGraph(P.LoadGraphRequest(csr=csr, positions=b"\x00" * 7, graph_revision=5))
    assert not r.ok and "multiple of 4" in r.error
    nan = np.zeros(12, np.float32)
    nan[4] = np.nan
    r = cl.LoadGraph(P.LoadGraphRequest(csr=csr, positions=_pos_bytes(nan), graph_revision=5))
    assert not r.ok and "non-finite" in r.error
    # no positions: the deterministic ring seed
    r = cl.LoadGraph(P.LoadGraphRequest(csr=csr, graph_revision=6))
    assert r.ok and r.n_nodes == 4
    assert np.array_equal(state.positions, W.ring_seed(4)) and state.frame == 0 and state.needs_engine_init


def test_health_reports_graph_and_engine_state(served):
    state, cl = served
    h = cl.Health(P.HealthRequest())
    assert h.ok and h.n_nodes == 3 and h.frame == 0 and h.graph_revision == 0
    assert h.requested_layout_id == "" and h.effective_layout_id == "" and h.fallback_reason == ""
    cl.LoadGraph(P.LoadGraphRequest(csr=jc.csr_to_bin(*_cycle4()), graph_revision=9))
    h = cl.Health(P.HealthRequest())
    assert h.n_nodes == 4 and h.graph_revision == 9


def test_list_engines(served):
    """list_engines_grpc.rs: every registered engine with a name, a description
    and the frozen kind string; the default id round-trips."""
    _, cl = served
    r = cl.ListEngines(P.ListEnginesRequest())
    assert r.default_id == "fa2-brute-hip"
    by_id = {e.id: e for e in r.engines}
    assert set(by_id) == {"fa2-bh-hip", "fa2-brute-hip", "sgd-stress-hip", "multilevel-hip"}
    for e in r.engines:
        assert e.display_name and e.description and e.kind in ("Physics", "Static")


def _code(call):
    with pytest.raises(grpc.RpcError) as ei:
        next(iter(call))
    return ei.value.code(), ei.value.details()


def test_subscribe_rejections_answer_before_any_engine(served):
    """service.rs subscribe: revision and generation preconditions, unknown
    layout ids and misaligned attributes are rejected before engine init."""
    state, cl = served
    cl.LoadGraph(P.LoadGraphRequest(csr=jc.csr_to_bin(*_cycle4()), graph_revision=42))
    code, msg = _code(cl.Subscribe(P.SubscribeRequest(graph_id="g", graph_revision=41)))
    assert code == grpc.StatusCode.FAILED_PRECONDITION and "graph revision mismatch" in msg
    code, msg = _code(cl.Subscribe(P.SubscribeRequest(graph_id="g", graph_revision=42, layout_id="cpu-spring")))
    assert code == grpc.StatusCode.INVALID_ARGUMENT and "unknown layout_id" in msg
    bad = P.SubscribeRequest(graph_id="g", layout_id="fa2-bh")
    bad.attributes.node_mass = b"\x00" * 6
    code, msg = _code(cl.Subscribe(bad))
    assert code == grpc.StatusCode.INVALID_ARGUMENT and "node_mass misaligned" in msg
    state.selection_generation = 5
    code, msg = _code(cl.Subscribe(P.SubscribeRequest(graph_id="g")))
    assert code == grpc.StatusCode.FAILED_PRECONDITION and "selection generation required" in msg
    code, msg = _code(cl.Subscribe(P.SubscribeRequest(graph_id="g", selection_generation=4)))
    assert code == grpc.StatusCode.FAILED_PRECONDITION and "stale selection generation" in msg


def test_reference_engine_ids_resolve_to_the_hip_engines():
    assert W.resolve("") == W.resolve("fa2-brute") == "fa2-brute-hip"
    assert W.resolve("fa2-bh") == "fa2-bh-hip" and W.resolve("multilevel") == "multilevel-hip"
    assert W.resolve("sgd-stress") == W.resolve("sgd-stress-gpu") == "sgd-stress-hip"
    assert W.resolve("cpu-spring") is None


def test_unimplemented_streams(served):
    _, cl = served
    with pytest.raises(grpc.RpcError) as ei:
        list(cl.ExchangeHalo(iter([P.HaloDelta(frame=1)])))
    assert ei.value.code() == grpc.StatusCode.UNIMPLEMENTED


def test_path_graph_is_the_reference_path():
    off, nb = W.path_graph(4)
    assert off.tolist() == [0, 1, 3, 5, 6] and nb.tolist() == [1, 0, 2, 1, 3, 2]


@pytest.mark.skipif(jc.device_count() > 0, reason="engine init succeeds where a GPU is present")
def test_failed_engine_init_is_an_error_not_a_cpu_fallback(served):
    """Without a GPU the HIP engine cannot initialise: the Subscribe fails with
    INTERNAL and Health reports the requested engine with no effective one
    (the reference would fall back to its CPU spring engine)."""
    _, cl = served
    code, msg = _code(cl.Subscribe(P.SubscribeRequest(graph_id="g", layout_id="fa2-bh")))
    assert code == grpc.StatusCode.INTERNAL and "engine init failed" in msg
    h = cl.Health(P.HealthRequest())
    assert h.requested_layout_id == "fa2-bh-hip" and h.effective_layout_id == "" and h.fallback_reason == ""
