"""The graph_compute.Compute worker streaming frames from the HIP engines: the
reference's grpc_stream.rs cases (a frame arrives, positions advance, a
same-size reload is honoured by revision) plus bit-exactness of the frames
against the engine run standalone, engine selection through Struct params and
the no-reset reconnect of the same selection generation."""
import time

import grpc
import numpy as np
import pytest

import jumpcannon as jc
from jumpcannon import compute_proto as P
from jumpcannon import worker as W

pytestmark = pytest.mark.gpu


def _serve(state, tick_hz=60.0):
    server, port, stop = W.serve(state, "127.0.0.1:0", tick_hz=tick_hz, max_workers=4)
    ch = grpc.insecure_channel(f"127.0.0.1:{port}")
    return server, stop, ch, W.ComputeClient(ch)


def _close(server, stop, ch, state):
    ch.close()
    stop.set()
    server.stop(0.5)
    state.close()


def _frames(stream, k, timeout=5.0):
    out, t0 = [], time.monotonic()
    for f in stream:
        out.append(f)
        if len(out) >= k or time.monotonic() - t0 > timeout:
            break
    stream.cancel()
    return out


def test_delivers_frames_that_advance_and_equal_the_standalone_engine():
    off, nb = W.path_graph(64)
    state = W.SimState(off, nb)
    assert state.init_engine("") == "fa2-brute-hip"
    server, stop, ch, cl = _serve(state)
    try:
        frames = _frames(cl.Subscribe(P.SubscribeRequest(graph_id="test")), 30)
    finally:
        _close(server, stop, ch, state)
    assert len(frames) == 30
    first, last = frames[0], frames[-1]
    assert first.n_nodes == 64 and len(first.positions) == 64 * 3 * 4 and first.frame >= 1
    assert [f.frame for f in frames] == list(range(first.frame, first.frame + 30))
    a = np.frombuffer(first.positions, "<f4")
    b = np.frombuffer(last.positions, "<f4")
    assert np.sqrt(((a - b) ** 2).sum()) > 0.0
    # frame k is the engine's k-th step from the ring seed
    e = jc.Fa2BruteEngine()
    e.init(off, nb, W.ring_seed(64))
    pos = e.step(int(first.frame))
    assert np.array_equal(pos, a)
    pos = e.step(int(last.frame - first.frame))
    assert np.array_equal(pos, b)


def test_same_cardinality_reload_rejects_stale_revision_and_stamps_new_frames():
    state = W.SimState(*W.path_graph(4))
    server, stop, ch, cl = _serve(state)
    try:
        r = cl.LoadGraph(P.LoadGraphRequest(csr=jc.csr_to_bin(*W.path_graph(4)), graph_revision=41))
        assert r.ok and r.graph_revision == 41
        seed = np.array([-3, 0, 0, 0, 3, 0, 3, 0, 0, 0, -3, 0], "<f4")
        cyc = (np.array([0, 2, 4, 6, 8], np.uint32), np.array([1, 3, 0, 2, 1, 3, 0, 2], np.uint32))
        r = cl.LoadGraph(P.LoadGraphRequest(csr=jc.csr_to_bin(*cyc), positions=seed.tobytes(), graph_revision=42))
        assert r.ok and r.n_nodes == 4 and r.graph_revision == 42
        with pytest.raises(grpc.RpcError) as ei:
            next(iter(cl.Subscribe(P.SubscribeRequest(graph_id="stale", graph_revision=41, layout_id="fa2-bh"))))
        assert ei.value.code() == grpc.StatusCode.FAILED_PRECONDITION
        frames = _frames(cl.Subscribe(P.SubscribeRequest(graph_id="new", graph_revision=42, layout_id="fa2-bh")), 2)
        assert frames and frames[0].graph_revision == 42 and frames[0].n_nodes == 4
        h = cl.Health(P.HealthRequest())
        assert h.requested_layout_id == h.effective_layout_id == "fa2-bh-hip" and h.graph_revision == 42
        # the first frame is one step of the engine from the supplied seed
        e = jc.Fa2BhEngine()
        e.init(*cyc, seed)
        assert np.array_equal(np.frombuffer(frames[0].positions, "<f4"), e.step(int(frames[0].frame)))
    finally:
        _close(server, stop, ch, state)


def test_engine_selection_with_struct_params_and_no_reset_reconnect():
    off, nb = jc.ba_graph(500, 3, 11)
    state = W.SimState(off, nb)
    server, stop, ch, cl = _serve(state)
    try:
        req = P.SubscribeRequest(graph_id="g", layout_id="sgd-stress", selection_generation=1)
        req.params.fields["n_pivots"].number_value = 20.0
        f1 = _frames(cl.Subscribe(req), 3)
        assert len(f1) == 3 and f1[0].n_nodes == 500 and f1[0].selection_generation == 1
        h = cl.Health(P.HealthRequest())
        assert h.effective_layout_id == "sgd-stress-hip" and h.selection_generation == 1
        # the same generation again: the running engine is kept, frames continue
        f2 = _frames(cl.Subscribe(P.SubscribeRequest(graph_id="g", layout_id="sgd-stress", selection_generation=1)), 2)
        assert f2[0].frame > f1[-1].frame
        # a newer generation selects again
        f3 = _frames(cl.Subscribe(P.SubscribeRequest(graph_id="g", layout_id="fa2-bh", selection_generation=2)), 2)
        assert f3[0].selection_generation == 2
        assert cl.Health(P.HealthRequest()).effective_layout_id == "fa2-bh-hip"
    finally:
        _close(server, stop, ch, state)
