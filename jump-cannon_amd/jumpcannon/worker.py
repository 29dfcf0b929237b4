"""The graph-compute worker over the HIP engines: the reference's gRPC service
(crates/graph-compute/src/service.rs:168-350), its simulation tick
(service.rs:642-736), its shared state (sim.rs:385-648) and its entry point
(src/main.rs), served with grpcio over the wire contract of compute_proto.

    python -m jumpcannon.worker     # GRAPH_COMPUTE_GRAPH_PATH, _LAYOUT_ID, _ADDR, _TICK_HZ

Engines: "fa2-brute-hip" (the default, as "fa2-brute" is the reference
registry's), "fa2-bh-hip", "sgd-stress-hip" and "multilevel-hip".  A Subscribe
may also name the reference engine each one replaces ("fa2-brute", "fa2-bh",
"sgd-stress", "sgd-stress-gpu", "multilevel"); other ids (the CPU spring and
geometric engines) are unknown.  Where the reference falls back to its CPU spring engine when
a GPU engine fails to initialise (sim.rs:607-642), this worker fails the
Subscribe with INTERNAL and reports the failure in Health: there is no CPU
path.  TopoFisheye and ExchangeHalo answer UNIMPLEMENTED (the reference's
answer for ExchangeHalo without a halo provider, service.rs:382-411).
"""
from __future__ import annotations

import concurrent.futures as cf
import math
import os
import queue
import sys
import threading
import time
from dataclasses import dataclass

import grpc
import numpy as np
from google.protobuf import json_format

from . import (Fa2BhEngine, Fa2BruteEngine, JcError, LayoutEngine, MultilevelEngine, SgdStressEngine, _ptr, csr_from_bin,
               csr_from_pairs, lib, load_csr_bin)
from . import compute_proto as P

# the registry default is the brute-force engine (engines/mod.rs:579)
DEFAULT_ID = "fa2-brute-hip"
ENGINES = {"fa2-bh-hip": Fa2BhEngine, "fa2-brute-hip": Fa2BruteEngine, "sgd-stress-hip": SgdStressEngine,
           "multilevel-hip": MultilevelEngine}
ALIASES = {"fa2-bh": "fa2-bh-hip", "fa2-brute": "fa2-brute-hip", "sgd-stress": "sgd-stress-hip",
           "sgd-stress-gpu": "sgd-stress-hip", "multilevel": "multilevel-hip"}
BROADCAST_DEPTH = 32  # frames a subscriber may lag before it drops the oldest (sim.rs:425-427)


def resolve(layout_id: str) -> str | None:
    """Registry id for a Subscribe's layout_id ("" is the default); None if unknown."""
    if not layout_id:
        return DEFAULT_ID
    if layout_id in ENGINES:
        return layout_id
    return ALIASES.get(layout_id)


def ring_seed(n: int) -> np.ndarray:
    """The deterministic seed of a graph loaded without positions (sim.rs:371-384)."""
    out = np.empty(3 * n, np.float32)
    if n:
        lib().jc_gen_ring_seed(n, _ptr(out))
    return out


def path_graph(n: int):
    """CsrGraph::path(n) (sim.rs:33-52): node i linked to i-1 and i+1."""
    pairs = np.stack([np.arange(n - 1, dtype=np.uint32), np.arange(1, n, dtype=np.uint32)], 1) if n > 1 else \
        np.zeros((0, 2), np.uint32)
    return csr_from_pairs(n, pairs.reshape(-1))


def struct_to_params(s) -> dict:
    """google.protobuf.Struct -> engine settings (service.rs:547-612): every
    Struct number is a double; whole, in-range values become integers so that
    integer settings decode."""
    def conv(v):
        if isinstance(v, float) and math.isfinite(v) and v == int(v) and -2.0 ** 63 <= v <= 2.0 ** 64 - 1:
            return int(v)
        if isinstance(v, dict):
            return {k: conv(x) for k, x in v.items()}
        if isinstance(v, list):
            return [conv(x) for x in v]
        return v
    return conv(json_format.MessageToDict(s))


class RpcError(Exception):
    def __init__(self, code: grpc.StatusCode, message: str):
        super().__init__(message)
        self.code = code
        self.message = message


@dataclass
class EngineStatus:
    """Health's view of the engine (sim.rs:356-363)."""
    requested_layout_id: str = ""
    effective_layout_id: str = ""
    fallback_reason: str = ""


@dataclass
class ActiveEngine:
    engine: LayoutEngine
    layout_id: str
    graph_revision: int
    selection_generation: int


def _attributes(a) -> None:
    """proto_attrs_to_host (service.rs:129-166): each present field is raw LE
    u32/f32, so its length must be a multiple of 4."""
    for name in ("node_class", "node_coordination", "node_mass", "edge_len"):
        b = getattr(a, name)
        if len(b) % 4:
            raise RpcError(grpc.StatusCode.INVALID_ARGUMENT,
                           f"GraphAttributes.{name} misaligned: len {len(b)} is not a multiple of 4")


class SimState:
    """The state the tick loop and the service share (sim.rs:385-648)."""

    def __init__(self, offsets, neighbors):
        self.offsets = np.asarray(offsets, np.uint32)
        self.neighbors = np.asarray(neighbors, np.uint32)
        self.graph_revision = 0
        self.selection_generation = 0
        self.needs_engine_init = True
        self.status = EngineStatus()
        self.positions = ring_seed(self.n_nodes)
        self.frame = 0
        self.active: ActiveEngine | None = None
        self._lock = threading.RLock()   # every field above
        self._swap = threading.Lock()    # graph replacement against engine initialisation
        self._subs: set[queue.Queue] = set()

    @property
    def n_nodes(self) -> int:
        return len(self.offsets) - 1

    # -- LoadGraph (sim.rs:449-487)
    def load_graph(self, offsets, neighbors, positions, graph_revision: int) -> int:
        n = len(offsets) - 1
        if positions is None:
            pos = ring_seed(n)
        else:
            pos = np.asarray(positions, np.float32)
            if len(pos) != 3 * n:
                raise ValueError(f"positions length {len(pos)} != 3 * n_nodes ({3 * n})")
            if not np.all(np.isfinite(pos)):
                raise ValueError("positions contain a non-finite component")
        with self._swap, self._lock:
            self.graph_revision = graph_revision
            self.offsets = np.asarray(offsets, np.uint32)
            self.neighbors = np.asarray(neighbors, np.uint32)
            self.positions = pos.copy()
            self.frame = 0
            self.active = None   # the next Subscribe re-inits on the new graph
            self.needs_engine_init = True
            self.status.effective_layout_id = ""
            self.status.fallback_reason = ""
        return n

    # -- engine selection (sim.rs:496-648)
    def init_engine(self, layout_id: str = "", params=None) -> str:
        return self.init_engine_for_generation(layout_id, params, self.selection_generation)

    def init_engine_for_generation(self, layout_id: str, params, selection_generation: int) -> str:
        with self._swap:
            with self._lock:
                current = self.selection_generation
                if selection_generation != 0 and selection_generation < current:
                    raise ValueError(f"stale selection generation {selection_generation}; worker has {current}")
                if selection_generation != 0 and selection_generation == current and not self.needs_engine_init:
                    if self.active is not None:
                        return self.active.layout_id
                    return self.status.effective_layout_id or DEFAULT_ID
                graph_revision = self.graph_revision
                off, nb = self.offsets, self.neighbors
                if len(self.positions) != 3 * self.n_nodes:
                    self.positions = ring_seed(self.n_nodes)   # self-heal a mismatched buffer
                pos = self.positions.copy()
            chosen = resolve(layout_id)
            if chosen is None:
                raise ValueError(f"unknown layout_id {layout_id!r}")
            with self._lock:
                self.status = EngineStatus(requested_layout_id=chosen)
            engine = ENGINES[chosen]()
            try:
                if params:
                    engine.set_params(params)
                engine.init(off, nb, pos)
            except JcError as e:
                engine.close()
                with self._lock:
                    self.status.fallback_reason = ""
                raise RuntimeError(f"engine {chosen} init failed: {e}") from e
            with self._lock:
                self.status.effective_layout_id = chosen
                self.selection_generation = selection_generation
                if self.active is not None and self.active.engine is not engine:
                    self.active.engine.close()
                self.active = ActiveEngine(engine, chosen, graph_revision, selection_generation)
                self.needs_engine_init = False
            return chosen

    # -- broadcast (a bounded queue per subscriber; a lagging one drops its oldest frames)
    def subscribe(self) -> queue.Queue:
        q: queue.Queue = queue.Queue(maxsize=BROADCAST_DEPTH)
        with self._lock:
            self._subs.add(q)
        return q

    def unsubscribe(self, q: queue.Queue) -> None:
        with self._lock:
            self._subs.discard(q)

    def _send(self, delta) -> None:
        with self._lock:
            subs = list(self._subs)
        for q in subs:
            while True:
                try:
                    q.put_nowait(delta)
                    break
                except queue.Full:
                    try:
                        q.get_nowait()
                    except queue.Empty:
                        pass

    # -- one tick of the simulation loop (service.rs:648-734); True if a frame went out
    def tick(self) -> bool:
        with self._lock:
            active, self.active = self.active, None   # taken out across the step
            if active is None:
                return False
            if active.graph_revision != self.graph_revision or \
                    active.selection_generation != self.selection_generation:
                active.engine.close()   # stale engine after a control-plane change
                return False
            if active.engine.is_halted():
                self.active = active    # converged: keep it and idle
                return False
        try:
            out = active.engine.step(1)
        except JcError as e:
            print(f"[graph-compute] engine step failed, dropping the engine: {e}", file=sys.stderr, flush=True)
            active.engine.close()
            with self._lock:
                if active.selection_generation == self.selection_generation:
                    self.needs_engine_init = True
            return False
        with self._lock:
            if active.graph_revision != self.graph_revision or \
                    active.selection_generation != self.selection_generation or \
                    len(out) != 3 * self.n_nodes:
                # a LoadGraph or a new selection landed during the step (or the
                # engine stepped a coarser level): drop the frame and the engine
                active.engine.close()
                return False
            if self.active is None:
                self.active = active
            else:
                active.engine.close()
            self.positions = out
            self.frame += 1
            delta = P.PositionDelta(frame=self.frame, positions=out.astype("<f4").tobytes(),
                                    n_nodes=len(out) // 3, graph_revision=self.graph_revision,
                                    selection_generation=self.selection_generation)
        self._send(delta)
        return True

    def close(self) -> None:
        with self._lock:
            if self.active is not None:
                self.active.engine.close()
                self.active = None


def run_sim_loop(state: SimState, tick_hz: float, stop: threading.Event) -> None:
    """Ticks at tick_hz, skipping missed ticks (service.rs:642-652)."""
    period = 1.0 / max(tick_hz, 1.0)
    nxt = time.monotonic()
    while not stop.is_set():
        state.tick()
        nxt += period
        now = time.monotonic()
        if nxt < now:
            nxt = now
        stop.wait(nxt - now)


class ComputeService:
    """graph_compute.Compute (service.rs:168-350)."""

    def __init__(self, state: SimState, stop: threading.Event | None = None):
        self.state = state
        self.stop = stop or threading.Event()

    def Subscribe(self, req, context):
        st = self.state
        try:
            with st._lock:
                active_revision = st.graph_revision
                current = st.selection_generation
                needs = st.needs_engine_init
            if req.graph_revision != 0 and req.graph_revision != active_revision:
                raise RpcError(grpc.StatusCode.FAILED_PRECONDITION,
                               f"graph revision mismatch: subscriber expects {req.graph_revision}, "
                               f"worker has {active_revision}")
            if req.selection_generation == 0 and current != 0:
                raise RpcError(grpc.StatusCode.FAILED_PRECONDITION,
                               f"selection generation required: worker has {current}")
            if req.selection_generation != 0 and req.selection_generation < current:
                raise RpcError(grpc.StatusCode.FAILED_PRECONDITION,
                               f"stale selection generation {req.selection_generation}: worker has {current}")
            has_params = req.HasField("params")
            want_select = needs or req.selection_generation > current or \
                (req.selection_generation == 0 and (bool(req.layout_id) or has_params))
            if want_select:
                if req.layout_id and resolve(req.layout_id) is None:
                    raise RpcError(grpc.StatusCode.INVALID_ARGUMENT, f"unknown layout_id {req.layout_id!r}")
                params = struct_to_params(req.params) if has_params else None
                if req.HasField("attributes"):
                    _attributes(req.attributes)
                try:
                    st.init_engine_for_generation(req.layout_id, params, req.selection_generation)
                except (RuntimeError, ValueError, JcError) as e:
                    raise RpcError(grpc.StatusCode.INTERNAL, f"engine init failed: {e}")
        except RpcError as e:
            context.abort(e.code, e.message)
        q = st.subscribe()   # before the first frame can go out
        try:
            while context.is_active() and not self.stop.is_set():
                try:
                    yield q.get(timeout=0.25)
                except queue.Empty:
                    continue
        finally:
            st.unsubscribe(q)

    def Health(self, req, context):
        st = self.state
        with st._lock:
            return P.HealthResponse(ok=True, n_nodes=st.n_nodes, frame=st.frame,
                                    requested_layout_id=st.status.requested_layout_id,
                                    effective_layout_id=st.status.effective_layout_id,
                                    fallback_reason=st.status.fallback_reason,
                                    graph_revision=st.graph_revision,
                                    selection_generation=st.selection_generation)

    def LoadGraph(self, req, context):
        st = self.state
        with st._lock:
            current = st.graph_revision
        try:
            off, nb = csr_from_bin(req.csr)
        except JcError as e:
            return P.LoadGraphResponse(ok=False, error=f"decode CSR: {e}", graph_revision=current)
        positions = None
        if req.positions:
            if len(req.positions) % 4:
                return P.LoadGraphResponse(ok=False, error="positions byte length not a multiple of 4",
                                           graph_revision=current)
            positions = np.frombuffer(req.positions, "<f4").astype(np.float32)
        try:
            n = st.load_graph(off, nb, positions, req.graph_revision)
        except ValueError as e:
            return P.LoadGraphResponse(ok=False, error=str(e), graph_revision=current)
        return P.LoadGraphResponse(ok=True, n_nodes=n, graph_revision=req.graph_revision)

    def ListEngines(self, req, context):
        out = []
        for eid, cls in ENGINES.items():
            e = cls()
            d = e.descriptor
            e.close()
            out.append(P.EngineDescriptor(id=d["id"], display_name=d["display_name"],
                                          description=d["description"],
                                          kind="Physics" if d["kind"] == "physics" else "Static"))
        return P.ListEnginesResponse(engines=out, default_id=DEFAULT_ID)

    def TopoFisheye(self, req_iter, context):
        context.abort(grpc.StatusCode.UNIMPLEMENTED, "TopoFisheye is not served by this worker")

    def ExchangeHalo(self, req_iter, context):
        context.abort(grpc.StatusCode.UNIMPLEMENTED, "ExchangeHalo needs a halo provider")


def _handler(svc: ComputeService) -> grpc.GenericRpcHandler:
    handlers = {}
    for rpc, (req, resp, cs, ss) in P.METHODS.items():
        fn = getattr(svc, rpc)
        de = P.message(req).FromString
        se = P.message(resp).SerializeToString
        make = {(False, False): grpc.unary_unary_rpc_method_handler,
                (False, True): grpc.unary_stream_rpc_method_handler,
                (True, True): grpc.stream_stream_rpc_method_handler}[(cs, ss)]
        handlers[rpc] = make(fn, request_deserializer=de, response_serializer=se)
    return grpc.method_handlers_generic_handler(P.SERVICE, handlers)


def serve(state: SimState, addr: str = "[::]:50051", tick_hz: float = 30.0, max_workers: int = 16,
          start_loop: bool = True):
    """Starts the gRPC server and the tick loop.  Returns (server, bound port,
    stop event); set the event and call server.stop() to shut down."""
    stop = threading.Event()
    svc = ComputeService(state, stop)
    server = grpc.server(cf.ThreadPoolExecutor(max_workers=max_workers))
    server.add_generic_rpc_handlers((_handler(svc),))
    port = server.add_insecure_port(addr)
    if port == 0:
        raise RuntimeError(f"cannot bind {addr}")
    server.start()
    if start_loop:
        threading.Thread(target=run_sim_loop, args=(state, tick_hz, stop), daemon=True,
                         name="graph-compute-tick").start()
    return server, port, stop


class ComputeClient:
    """A client of graph_compute.Compute over a grpcio channel."""

    def __init__(self, channel: grpc.Channel):
        for rpc, (req, resp, cs, ss) in P.METHODS.items():
            make = {(False, False): channel.unary_unary, (False, True): channel.unary_stream,
                    (True, True): channel.stream_stream}[(cs, ss)]
            setattr(self, rpc, make(P.method_path(rpc), request_serializer=P.message(req).SerializeToString,
                                    response_deserializer=P.message(resp).FromString))


def main() -> int:
    """src/main.rs: the graph from GRAPH_COMPUTE_GRAPH_PATH (binary CSR) or a
    1,024-node path, the engine from GRAPH_COMPUTE_LAYOUT_ID, the address from
    GRAPH_COMPUTE_ADDR ([::]:50051) and the tick rate from GRAPH_COMPUTE_TICK_HZ (30)."""
    path = os.environ.get("GRAPH_COMPUTE_GRAPH_PATH", "")
    if path:
        off, nb = load_csr_bin(path)
        source = f"file: {path} ({len(off) - 1} nodes, {len(nb)} edges)"
    else:
        off, nb = path_graph(1024)
        source = "path-graph (1024 nodes)"
    print(f"[graph-compute] graph source: {source}", flush=True)
    state = SimState(off, nb)
    try:
        eid = state.init_engine(os.environ.get("GRAPH_COMPUTE_LAYOUT_ID", ""))
        print(f"[graph-compute] layout engine initialized: {eid}", flush=True)
    except (RuntimeError, ValueError, JcError) as e:
        print(f"[graph-compute] failed to initialize layout engine: {e}", file=sys.stderr, flush=True)
    try:
        tick_hz = float(os.environ.get("GRAPH_COMPUTE_TICK_HZ", "30"))
    except ValueError:
        tick_hz = 30.0
    addr = os.environ.get("GRAPH_COMPUTE_ADDR", "[::]:50051")
    server, port, stop = serve(state, addr, tick_hz)
    print(f"[graph-compute] serving on {addr} (port {port}), tick {tick_hz} Hz, {state.n_nodes} nodes",
          flush=True)
    try:
        server.wait_for_termination()
    except KeyboardInterrupt:
        pass
    finally:
        stop.set()
        server.stop(1.0)
        state.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
