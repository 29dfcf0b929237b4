"""Python front-end of libjc_hip.so (the MI355X layout / graph-compute engines).

Mirrors the reference's engine API (crates/graph-compute/src/engines/mod.rs:383-444):
``set_params`` -> ``init`` -> ``step`` per tick, plus the analytics free
functions (``pagerank``, ``spmv`` / ``spmv_f16`` / ``spmv_hybrid``, ``bfs``,
``connected_components``).  Everything runs through the C ABI declared in
include/jc.h; there is no CPU fallback -- a missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(PKG_DIR), "libjc_hip.so")

JC_OK = 0
JC_E_NODEVICE = -1
JC_E_INVALID = -2
JC_E_OOM = -3
JC_E_HIP = -4
JC_E_STATE = -5
JC_E_DANGLING = -6

_ERRNAMES = {JC_E_NODEVICE: "NODEVICE", JC_E_INVALID: "INVALID", JC_E_OOM: "OOM",
             JC_E_HIP: "HIP", JC_E_STATE: "STATE", JC_E_DANGLING: "DANGLING"}


class JcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class NoDeviceError(JcError):
    pass


class Fa2BhSettings(C.Structure):
    """Field-for-field Fa2BhSettings (crates/graph-compute/src/engines/fa2_bh.rs:85-125)."""

    _fields_ = [("gravity", C.c_float), ("strong_gravity", C.c_uint32),
                ("scaling_ratio", C.c_float), ("edge_weight_influence", C.c_float),
                ("jitter_tolerance", C.c_float), ("lin_log_mode", C.c_uint32),
                ("prevent_overlap", C.c_uint32), ("time_step", C.c_float),
                ("theta", C.c_float), ("max_displacement", C.c_float)]


class Fa2Settings(C.Structure):
    """Field-for-field Fa2Settings of the brute-force engine (fa2_brute.rs:63-93)."""

    _fields_ = [("gravity", C.c_float), ("strong_gravity", C.c_uint32),
                ("scaling_ratio", C.c_float), ("edge_weight_influence", C.c_float),
                ("jitter_tolerance", C.c_float), ("lin_log_mode", C.c_uint32),
                ("prevent_overlap", C.c_uint32), ("time_step", C.c_float),
                ("max_displacement", C.c_float)]


class SgdSettings(C.Structure):
    """Field-for-field SgdStressSettings (crates/graph-compute/src/engines/sgd_stress.rs:59-98)."""

    _fields_ = [("n_pivots", C.c_uint32), ("sweeps_per_step", C.c_uint32),
                ("eta_max", C.c_float), ("eta_min", C.c_float),
                ("n_anneal_steps", C.c_uint32), ("reserved", C.c_uint32), ("seed", C.c_uint64)]


class Descriptor(C.Structure):
    _fields_ = [("id", C.c_char_p), ("kind", C.c_char_p), ("display_name", C.c_char_p),
                ("description", C.c_char_p), ("needs_edges", C.c_uint32),
                ("needs_cpu_positions", C.c_uint32), ("needs_gpu_positions_buffer", C.c_uint32)]


_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_vp = C.c_void_p

# Symbol -> (restype, argtypes).  Pointer arguments are c_void_p so that None
# (NULL) and numpy arrays (via .ctypes.data) are both accepted.
SIGNATURES = {
    "jc_last_error": (C.c_char_p, []),
    "jc_device_count": (C.c_int, []),
    "jc_fa2bh_default_settings": (None, [_vp]),
    "jc_fa2bh_create": (C.c_int, [_vp, _vp]),
    "jc_fa2_default_settings": (None, [_vp]),
    "jc_fa2brute_create": (C.c_int, [_vp, _vp]),
    "jc_sgd_default_settings": (None, [_vp]),
    "jc_sgd_create": (C.c_int, [_vp, _vp]),
    "jc_engine_descriptor": (C.c_int, [_vp, _vp]),
    "jc_engine_set_params_json": (C.c_int, [_vp, C.c_char_p, C.c_size_t]),
    "jc_engine_init": (C.c_int, [_vp, C.c_uint32, _vp, _vp, C.c_uint64, _vp]),
    "jc_engine_reinit": (C.c_int, [_vp, C.c_uint32, _vp, _vp, C.c_uint64, _vp]),
    "jc_engine_step": (C.c_int, [_vp, C.c_uint32, _vp]),
    "jc_engine_sync": (C.c_int, [_vp]),
    "jc_engine_get_state": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "jc_engine_set_state": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint64]),
    "jc_engine_is_halted": (C.c_int, [_vp]),
    "jc_engine_destroy": (None, [_vp]),
    "jc_fa2bh_forces_at": (C.c_int, [_vp, _vp, _vp, _vp]),
    "jc_fa2bh_tree": (C.c_int, [_vp, C.c_uint32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "jc_fa2bh_step_timed": (C.c_int, [_vp, C.c_uint32, _vp, _vp]),
    "jc_fa2bh_profile": (C.c_int, [_vp, C.c_uint32, _vp, C.c_uint32]),
    "jc_fa2bh_last_stats": (C.c_int, [_vp, _vp, _vp, _vp]),
    "jc_fa2bh_walk_mode": (C.c_int, [_vp, C.c_int, _vp]),
    "jc_fa2bh_phase_name": (C.c_char_p, [C.c_uint32]),
    "jc_fa2bh_phase_count": (C.c_uint32, []),
    "jc_sgd_pivots": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "jc_sgd_time": (C.c_int, [_vp, C.c_uint32, _vp]),
    "jc_stress_plan_create": (C.c_int, [C.c_uint32, C.c_uint32, _vp, _vp, _vp]),
    "jc_stress_plan_eval": (C.c_int, [_vp, _vp, _vp]),
    "jc_stress_plan_destroy": (None, [_vp]),
    "jc_pagerank": (C.c_int, [C.c_uint32, _vp, _vp, C.c_uint64, C.c_float, C.c_uint32,
                              C.c_double, _vp, _vp, _vp]),
    "jc_pagerank_plan_create": (C.c_int, [C.c_uint32, _vp, _vp, C.c_uint64, C.c_float, _vp]),
    "jc_pagerank_plan_run": (C.c_int, [_vp, C.c_uint32, C.c_double, _vp, _vp, _vp, _vp]),
    "jc_pagerank_plan_time": (C.c_int, [_vp, C.c_uint32, _vp]),
    "jc_pagerank_plan_destroy": (None, [_vp]),
    "jc_pagerank_noalloc": (C.c_int, [C.c_uint32, _vp, _vp, C.c_uint64, C.c_float, C.c_uint32, C.c_double,
                                      _vp, _vp, _vp, _vp]),
    "jc_spmv": (C.c_int, [C.c_uint32, _vp, _vp, _vp, C.c_uint64, _vp, _vp]),
    "jc_spmv_f16": (C.c_int, [C.c_uint32, _vp, _vp, _vp, C.c_uint64, _vp, _vp]),
    "jc_spmv_hybrid": (C.c_int, [C.c_uint32, _vp, _vp, _vp, C.c_uint64, _vp, _vp]),
    "jc_spmv_plan_create": (C.c_int, [C.c_uint32, _vp, _vp, _vp, C.c_uint64, C.c_int, _vp]),
    "jc_spmv_plan_run": (C.c_int, [_vp, _vp, _vp, _vp]),
    "jc_spmv_plan_time": (C.c_int, [_vp, C.c_uint32, _vp]),
    "jc_spmv_plan_destroy": (None, [_vp]),
    "jc_multilevel_create": (C.c_int, [_vp]),
    "jc_multilevel_state": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "jc_bfs": (C.c_int, [C.c_uint32, _vp, _vp, C.c_uint64, C.c_uint32, _vp]),
    "jc_bfs_plan_create": (C.c_int, [C.c_uint32, _vp, _vp, C.c_uint64, _vp]),
    "jc_bfs_plan_run": (C.c_int, [_vp, C.c_uint32, _vp, _vp]),
    "jc_bfs_plan_destroy": (None, [_vp]),
    "jc_connected_components": (C.c_int, [C.c_uint32, _vp, _vp, C.c_uint64, _vp]),
    "jc_largest_component": (C.c_int, [C.c_uint32, _vp, _vp, C.c_uint64, _vp, _vp, _vp, _vp, _vp]),
    "jc_fa2bh_batch_create": (C.c_int, [_vp, _vp]),
    "jc_fa2bh_batch_init": (C.c_int, [_vp, C.c_uint32, _vp, _vp, _vp, _vp, _vp]),
    "jc_fa2bh_batch_step": (C.c_int, [_vp, C.c_uint32, _vp]),
    "jc_fa2bh_batch_speeds": (C.c_int, [_vp, _vp]),
    "jc_fa2bh_batch_step_timed": (C.c_int, [_vp, C.c_uint32, _vp, _vp]),
    "jc_fa2bh_batch_set_schedule": (C.c_int, [_vp, C.c_uint32, C.c_uint32]),
    "jc_fa2bh_batch_schedule": (C.c_int, [_vp, _vp]),
    "jc_fa2bh_batch_set_tuning": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "jc_fa2bh_batch_destroy": (None, [_vp]),
    "jc_mpnn_param_count": (C.c_uint64, []),
    "jc_mpnn_create": (C.c_int, [C.c_uint64, _vp]),
    "jc_mpnn_destroy": (None, [_vp]),
    "jc_mpnn_get_params": (C.c_int, [_vp, _vp]),
    "jc_mpnn_set_params": (C.c_int, [_vp, _vp]),
    "jc_mpnn_get_grads": (C.c_int, [_vp, _vp]),
    "jc_gen_ring_seed": (None, [C.c_uint32, _vp]),
    "jc_mpnn_get_adam": (C.c_int, [_vp, _vp, _vp, _vp]),
    "jc_mpnn_set_adam": (C.c_int, [_vp, _vp, _vp, C.c_uint64]),
    "jc_mpnn_set_graph": (C.c_int, [_vp, C.c_uint32, _vp, _vp, C.c_uint64, C.c_uint64, C.c_uint32, _vp]),
    "jc_mpnn_features": (C.c_int, [_vp, _vp]),
    "jc_mpnn_activations": (C.c_int, [_vp, C.c_int, _vp]),
    "jc_mpnn_forward": (C.c_int, [_vp, _vp, _vp]),
    "jc_mpnn_forward_scaled": (C.c_int, [_vp, C.c_double, _vp, _vp]),
    "jc_mpnn_set_target": (C.c_int, [_vp, _vp, C.c_uint32, C.c_uint64]),
    "jc_mpnn_loss_grad": (C.c_int, [_vp, _vp]),
    "jc_mpnn_adam": (C.c_int, [_vp, C.c_float]),
    "jc_mpnn_train": (C.c_int, [_vp, C.c_uint32, C.c_float, _vp]),
    "jc_gen_ba_pairs": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint64, _vp, C.c_uint64]),
    "jc_gen_sbm_pairs": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                      C.c_uint64, _vp, C.c_uint64]),
    "jc_gen_rgg_pairs": (C.c_uint64, [C.c_uint32, C.c_double, C.c_uint64, _vp, _vp, C.c_uint64]),
    "jc_csr_from_pairs": (C.c_uint64, [C.c_uint32, _vp, C.c_uint64, _vp, _vp]),
    "jc_gen_uniform_cube": (None, [C.c_uint32, C.c_float, C.c_uint64, _vp]),
    "jc_gen_ball_seed": (None, [C.c_uint32, C.c_uint64, _vp]),
    "jc_step_stats": (C.c_int, [C.c_uint32, _vp, _vp, _vp]),
    "jc_csr_bin_header": (C.c_int, [_vp, C.c_size_t, _vp, _vp]),
    "jc_csr_from_bin": (C.c_int, [_vp, C.c_size_t, _vp, _vp]),
    "jc_csr_bin_size": (C.c_uint64, [C.c_uint32, C.c_uint64]),
    "jc_csr_to_bin": (C.c_int, [C.c_uint32, _vp, _vp, C.c_uint64, _vp, C.c_uint64]),
}

_lib = None


def lib():
    """Loads libjc_hip.so from the package tree; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python jump-cannon_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data


def _check(rc: int):
    if rc != JC_OK:
        msg = lib().jc_last_error().decode(errors="replace")
        if rc == JC_E_NODEVICE:
            raise NoDeviceError(rc, msg)
        raise JcError(rc, msg)


def device_count() -> int:
    return int(lib().jc_device_count())


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _need_len(a, want: int, what: str):
    """The C entry points read exactly `want` elements from this buffer: a
    shorter one would be read past its end, a longer one means a caller error."""
    if a is not None and len(a) != want:
        raise JcError(JC_E_INVALID, f"{what} length {len(a)} != {want}")
    return a


class LayoutEngine:
    """One engine handle (the reference LayoutEngine trait object)."""

    def __init__(self, handle):
        self._h = handle
        self.n = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().jc_engine_destroy(h)
            except Exception:
                pass
            self._h = None

    def close(self):
        self.__del__()

    @property
    def descriptor(self) -> dict:
        d = Descriptor()
        _check(lib().jc_engine_descriptor(self._h, C.addressof(d)))
        return {"id": d.id.decode(), "kind": d.kind.decode(),
                "display_name": d.display_name.decode(), "description": d.description.decode(),
                "needs_edges": bool(d.needs_edges),
                "needs_cpu_positions": bool(d.needs_cpu_positions),
                "needs_gpu_positions_buffer": bool(d.needs_gpu_positions_buffer)}

    def set_params(self, params) -> None:
        """LayoutEngine::set_params: a dict (or JSON text); None keeps the settings."""
        text = params if isinstance(params, str) else json.dumps(params)
        b = text.encode()
        _check(lib().jc_engine_set_params_json(self._h, b, len(b)))

    def init(self, offsets, neighbors, xyz) -> None:
        offsets = _u32(offsets)
        neighbors = _u32(neighbors)
        xyz = _f32(xyz)
        n = len(offsets) - 1
        if len(xyz) != 3 * n:
            raise JcError(JC_E_INVALID,
                          f"initial positions length {len(xyz)} != 3 * n_nodes {3 * n}")
        self._keep = (offsets, neighbors)
        _check(lib().jc_engine_init(self._h, n, _ptr(offsets), _ptr(neighbors) if len(neighbors) else None,
                                    len(neighbors), _ptr(xyz) if n else None))
        self.n = n

    def reinit(self, offsets, neighbors, xyz) -> None:
        offsets = _u32(offsets)
        neighbors = _u32(neighbors)
        xyz = _f32(xyz)
        n = len(offsets) - 1
        _need_len(xyz, 3 * n, "positions")
        _check(lib().jc_engine_reinit(self._h, n, _ptr(offsets), _ptr(neighbors) if len(neighbors) else None,
                                      len(neighbors), _ptr(xyz) if n else None))
        self.n = n

    def step(self, n_iters: int = 1, readback: bool = True):
        """LayoutEngine::step, n_iters times.  Returns positions (3n) or None."""
        out = np.empty(3 * self.n, np.float32) if readback else None
        _check(lib().jc_engine_step(self._h, n_iters, _ptr(out) if readback and self.n else None))
        return out

    def sync(self):
        _check(lib().jc_engine_sync(self._h))

    def get_state(self):
        xyz = np.empty(3 * self.n, np.float32)
        old = np.empty(3 * self.n, np.float32)
        se = np.empty(2, np.float32)
        it = np.zeros(1, np.uint64)
        _check(lib().jc_engine_get_state(self._h, _ptr(xyz), _ptr(old), _ptr(se), _ptr(it)))
        return {"positions": xyz, "old_force": old, "speed": float(se[0]),
                "speed_efficiency": float(se[1]), "iteration": int(it[0]), "speed_eff2": se}

    def set_state(self, positions=None, old_force=None, speed_eff2=None, iteration=0):
        p = _need_len(None if positions is None else _f32(positions), 3 * self.n, "positions")
        o = _need_len(None if old_force is None else _f32(old_force), 3 * self.n, "old_force")
        s = _need_len(None if speed_eff2 is None else _f32(speed_eff2), 2, "speed_eff2")
        _check(lib().jc_engine_set_state(self._h, _ptr(p), _ptr(o), _ptr(s), iteration))

    def is_halted(self) -> bool:
        return bool(lib().jc_engine_is_halted(self._h))


class MultilevelEngine(LayoutEngine):
    """Multilevel wrapper ("multilevel-hip"), the drop-in for
    crates/graph-compute/src/engines/multilevel.rs: coarsen -> solve the
    coarsest level -> prolong -> refine, with "fa2-brute-hip" (the default, as
    in multilevel.rs:176), "fa2-bh-hip" or "sgd-stress-hip" inside.  Settings are MultilevelSettings' keys, given as
    keyword arguments or through set_params."""

    ID = "multilevel-hip"

    def __init__(self, **settings):
        h = C.c_void_p()
        _check(lib().jc_multilevel_create(C.addressof(h)))
        super().__init__(h.value)
        if settings:
            self.set_params(settings)

    def state(self) -> dict:
        """The cascade position: level (0 = input graph), levels, the level's
        node count and the ticks taken at it."""
        v = np.zeros(4, np.uint32)
        _check(lib().jc_multilevel_state(self._h, _ptr(v[0:1]), _ptr(v[1:2]), _ptr(v[2:3]), _ptr(v[3:4])))
        return {"level": int(v[0]), "n_levels": int(v[1]), "n_level_nodes": int(v[2]),
                "sweeps_done": int(v[3])}

    def step(self, n_iters: int = 1, readback: bool = True):
        """n_iters ticks; returns the current level's positions (3 * n_level)."""
        out = np.empty(3 * max(self.n, 1), np.float32) if readback else None
        _check(lib().jc_engine_step(self._h, n_iters, _ptr(out) if readback and self.n else None))
        if not readback:
            return None
        return out[: 3 * self.state()["n_level_nodes"]]


def default_fa2bh_settings(**overrides) -> Fa2BhSettings:
    s = Fa2BhSettings()
    lib().jc_fa2bh_default_settings(C.addressof(s))
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


class Fa2BhEngine(LayoutEngine):
    """ForceAtlas2 + Barnes-Hut ("fa2-bh-hip"), the drop-in for fa2_bh.rs."""

    ID = "fa2-bh-hip"

    def __init__(self, settings: Fa2BhSettings | None = None, **overrides):
        s = settings if settings is not None else default_fa2bh_settings()
        for k, v in overrides.items():
            setattr(s, k, v)
        h = C.c_void_p()
        _check(lib().jc_fa2bh_create(C.addressof(s), C.addressof(h)))
        super().__init__(h.value)

    def forces_at(self, xyz):
        """One force pass at xyz: returns (F (3n) f32, [sum swing, sum traction] f64)."""
        xyz = _need_len(_f32(xyz), 3 * self.n, "positions")
        f = np.empty(3 * self.n, np.float32)
        sums = np.zeros(2, np.float64)
        _check(lib().jc_fa2bh_forces_at(self._h, _ptr(xyz), _ptr(f), _ptr(sums)))
        return f, sums

    def tree(self, cap: int | None = None):
        """The octree of the last tree build (pre-order arrays)."""
        n = self.n
        cap = cap if cap is not None else 23 * max(n, 1) + 8
        nt = np.zeros(1, np.uint32)
        nv = np.zeros(1, np.uint32)
        keys = np.zeros(max(n, 1), np.uint64)
        perm = np.zeros(max(n, 1), np.uint32)
        na = np.zeros(4 * cap, np.float32)
        info = np.zeros(cap, np.uint32)
        skip = np.zeros(cap, np.uint32)
        depth = np.zeros(cap, np.uint8)
        root4 = np.zeros(4, np.float32)
        _check(lib().jc_fa2bh_tree(self._h, cap, _ptr(nt), _ptr(nv), _ptr(keys), _ptr(perm),
                                   _ptr(na), _ptr(info), _ptr(skip), _ptr(depth), _ptr(root4)))
        N = int(nt[0])
        return {"n_tree": N, "n_valid": int(nv[0]), "keys": keys[:n], "perm": perm[:n],
                "node_a": na[: 4 * N].reshape(-1, 4), "info": info[:N], "skip": skip[:N],
                "depth": depth[:N], "root4": root4}

    def step_timed(self, n_iters: int):
        """Graph-replayed iterations with HIP events around each force kernel.
        Returns (sum of force-kernel ms, sum of whole-iteration ms)."""
        f = np.zeros(1, np.float64)
        it = np.zeros(1, np.float64)
        _check(lib().jc_fa2bh_step_timed(self._h, n_iters, _ptr(f), _ptr(it)))
        return float(f[0]), float(it[0])

    def walk_mode(self, mode: int | None = None) -> int:
        """Query (None) or force the walk: 0 per-lane rope, 1 wave-union, 2/3 two/four
        interleaved bodies per lane, 4/5/6 per-lane rope in 512/1024/64-thread
        blocks, -1 re-time and keep the fastest."""
        cur = np.zeros(1, np.int32)
        _check(lib().jc_fa2bh_walk_mode(self._h, -2 if mode is None else mode, _ptr(cur)))
        return int(cur[0])

    def last_stats(self) -> dict:
        nt = np.zeros(1, np.uint64)
        it = np.zeros(1, np.uint64)
        sums = np.zeros(2, np.float64)
        _check(lib().jc_fa2bh_last_stats(self._h, _ptr(nt), _ptr(it), _ptr(sums)))
        return {"n_tree": int(nt[0]), "interactions": int(it[0]), "sum_swing": float(sums[0]),
                "sum_traction": float(sums[1])}

    def profile(self, n_iters: int) -> dict:
        k = int(lib().jc_fa2bh_phase_count())
        out = np.zeros(k, np.float64)
        _check(lib().jc_fa2bh_profile(self._h, n_iters, _ptr(out), k))
        return {lib().jc_fa2bh_phase_name(i).decode(): float(out[i]) for i in range(k)}


def default_fa2_settings(**overrides) -> Fa2Settings:
    s = Fa2Settings()
    lib().jc_fa2_default_settings(C.addressof(s))
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


class Fa2BruteEngine(Fa2BhEngine):
    """Brute-force ForceAtlas2 ("fa2-brute-hip"), the drop-in for fa2_brute.rs:
    O(n^2) repulsion over every pair in index order.  The parity hooks are
    Fa2BhEngine's (forces_at, step_timed, last_stats, profile); it builds no
    octree, so tree() is empty.  walk_mode() picks its repulsion kernel: 0 one
    lane per body, 1 four, 2 sixteen; -1 (the default) times them at the first
    step and keeps the fastest; all give bit-identical forces."""

    ID = "fa2-brute-hip"

    def __init__(self, settings: Fa2Settings | None = None, **overrides):
        s = settings if settings is not None else default_fa2_settings()
        for k, v in overrides.items():
            setattr(s, k, v)
        h = C.c_void_p()
        _check(lib().jc_fa2brute_create(C.addressof(s), C.addressof(h)))
        LayoutEngine.__init__(self, h.value)


def default_sgd_settings(**overrides) -> SgdSettings:
    s = SgdSettings()
    lib().jc_sgd_default_settings(C.addressof(s))
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


class SgdStressEngine(LayoutEngine):
    """Sparse-pivot SGD stress ("sgd-stress-hip"), the drop-in for sgd_stress_gpu.rs."""

    ID = "sgd-stress-hip"

    def __init__(self, settings: SgdSettings | None = None, **overrides):
        s = settings if settings is not None else default_sgd_settings()
        for k, v in overrides.items():
            setattr(s, k, v)
        h = C.c_void_p()
        _check(lib().jc_sgd_create(C.addressof(s), C.addressof(h)))
        super().__init__(h.value)

    def pivots(self):
        """(pivots, node-major u16 distances, eta_max in force)."""
        k = np.zeros(1, np.uint32)
        em = np.zeros(1, np.float32)
        _check(lib().jc_sgd_pivots(self._h, _ptr(k), None, None, _ptr(em)))
        kk = int(k[0])
        piv = np.zeros(max(kk, 1), np.uint32)
        dist = np.zeros(max(kk * self.n, 1), np.uint16)
        _check(lib().jc_sgd_pivots(self._h, _ptr(k), _ptr(piv), _ptr(dist), _ptr(em)))
        return piv[:kk], dist[: kk * self.n], float(em[0])

    def time_steps(self, n_steps: int) -> float:
        ms = np.zeros(1, np.float64)
        _check(lib().jc_sgd_time(self._h, n_steps, _ptr(ms)))
        return float(ms[0])


class Fa2BhBatch:
    """Many independent graphs laid out together (one launch per phase for the
    whole batch); each graph evolves exactly as a standalone Fa2BhEngine."""

    def __init__(self, settings: Fa2BhSettings | None = None, **overrides):
        s = settings if settings is not None else default_fa2bh_settings()
        for k, v in overrides.items():
            setattr(s, k, v)
        h = C.c_void_p()
        _check(lib().jc_fa2bh_batch_create(C.addressof(s), C.addressof(h)))
        self._h = h.value
        self.sizes = None

    def init(self, graphs, positions):
        """graphs: list of (offsets, neighbors); positions: list of xyz arrays."""
        sizes = np.array([len(o) - 1 for o, _ in graphs], np.uint32)
        nnz = np.array([len(nb) for _, nb in graphs], np.uint64)
        off = np.concatenate([_u32(o) for o, _ in graphs])
        nbs = [_u32(nb) for _, nb in graphs]
        nb = np.concatenate(nbs) if sum(len(x) for x in nbs) else np.zeros(1, np.uint32)
        xyz = np.concatenate([_f32(p) for p in positions]) if len(positions) else np.zeros(0, np.float32)
        if len(xyz) != 3 * int(sizes.sum()):
            raise JcError(JC_E_INVALID, "positions do not match the graph sizes")
        self._keep = (sizes, nnz, off, nb)
        _check(lib().jc_fa2bh_batch_init(self._h, len(sizes), _ptr(sizes), _ptr(nnz), _ptr(off),
                                          _ptr(nb), _ptr(xyz) if len(xyz) else None))
        self.sizes = sizes

    def step(self, n_iters: int = 1, readback: bool = True):
        """Returns the list of per-graph positions (or None)."""
        tot = int(self.sizes.sum())
        out = np.empty(3 * max(tot, 1), np.float32) if readback else None
        _check(lib().jc_fa2bh_batch_step(self._h, n_iters, _ptr(out)))
        if not readback:
            return None
        parts, o = [], 0
        for n in self.sizes:
            parts.append(out[3 * o: 3 * (o + int(n))])
            o += int(n)
        return parts

    def speeds(self):
        se = np.zeros(2 * len(self.sizes), np.float32)
        _check(lib().jc_fa2bh_batch_speeds(self._h, _ptr(se)))
        return se.reshape(-1, 2)

    def step_timed(self, n_iters: int):
        ms = np.zeros(1, np.float64)
        bs = np.zeros(1, np.uint32)
        _check(lib().jc_fa2bh_batch_step_timed(self._h, n_iters, _ptr(ms), _ptr(bs)))
        return float(ms[0]), int(bs[0])

    SCHEDULES = {"auto": 0, "phased": 1, "persistent": 2}

    def set_schedule(self, mode: str = "auto", threads: int = 0):
        """'phased' (one launch per phase, hipGraph replay), 'persistent' (one
        launch per step, a workgroup per graph; threads 256/512/1024 or 0 for
        the tuner's pick) or 'auto' (the autotuner times both at the next step)."""
        if mode not in self.SCHEDULES:
            raise JcError(JC_E_INVALID, f"unknown schedule {mode!r}")
        _check(lib().jc_fa2bh_batch_set_schedule(self._h, self.SCHEDULES[mode], int(threads)))

    def set_tuning(self, force_block: int = 0, com_depth: int = 0, com_chunk: int = 0,
                   sort_key_bits: int = 0):
        """Fixes knobs of the phased schedule (0: the autotuner picks): force block
        size, centre-of-mass fusion depth and chunk, composite sort key width."""
        _check(lib().jc_fa2bh_batch_set_tuning(self._h, int(force_block), int(com_depth),
                                               int(com_chunk), int(sort_key_bits)))

    def schedule(self):
        """(mode, persistent workgroup size, force block size, centre-of-mass
        fusion depth, centre-of-mass chunk, sort key width)."""
        v = np.zeros(6, np.uint32)
        _check(lib().jc_fa2bh_batch_schedule(self._h, _ptr(v)))
        names = {v2: k for k, v2 in self.SCHEDULES.items()}
        return (names[int(v[0])],) + tuple(int(x) for x in v[1:])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().jc_fa2bh_batch_destroy(h)
            self._h = None


class StressJudge:
    """Scale-normalized stress (metrics.rs:31-59) over pivot terms, on the GPU."""

    def __init__(self, n: int, pivots, dist_node_major):
        self._piv = _u32(pivots)
        self._dist = np.ascontiguousarray(dist_node_major, np.uint16)
        self.n = int(n)
        _need_len(self._dist, self.n * len(self._piv), "pivot distance table")
        h = C.c_void_p()
        _check(lib().jc_stress_plan_create(n, len(self._piv), _ptr(self._piv), _ptr(self._dist),
                                           C.addressof(h)))
        self._h = h.value

    def __call__(self, xyz) -> float:
        x = _need_len(_f32(xyz), 3 * self.n, "positions")
        out = np.zeros(1, np.float64)
        _check(lib().jc_stress_plan_eval(self._h, _ptr(x), _ptr(out)))
        return float(out[0])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().jc_stress_plan_destroy(h)
            self._h = None


class Mpnn:
    """The warm-start network (3 message-passing layers, hidden 128, BF16 MFMA)."""

    def __init__(self, seed: int = 1):
        h = C.c_void_p()
        _check(lib().jc_mpnn_create(seed, C.addressof(h)))
        self._h = h.value
        self.n = 0
        self.P = int(lib().jc_mpnn_param_count())

    def params(self):
        out = np.empty(self.P, np.float32)
        _check(lib().jc_mpnn_get_params(self._h, _ptr(out)))
        return out

    def set_params(self, p):
        p = _need_len(_f32(p), self.P, "parameters")
        _check(lib().jc_mpnn_set_params(self._h, _ptr(p)))

    def grads(self):
        out = np.empty(self.P, np.float32)
        _check(lib().jc_mpnn_get_grads(self._h, _ptr(out)))
        return out

    def adam_state(self):
        """(first moments, second moments, step count) of the Adam optimizer."""
        m1 = np.empty(self.P, np.float32)
        m2 = np.empty(self.P, np.float32)
        t = np.zeros(1, np.uint64)
        _check(lib().jc_mpnn_get_adam(self._h, _ptr(m1), _ptr(m2), _ptr(t)))
        return m1, m2, int(t[0])

    def set_adam_state(self, m1, m2, step: int):
        m1 = _need_len(_f32(m1), self.P, "Adam first moments")
        m2 = _need_len(_f32(m2), self.P, "Adam second moments")
        _check(lib().jc_mpnn_set_adam(self._h, _ptr(m1), _ptr(m2), int(step)))

    def set_graph(self, offsets, neighbors, feat_seed: int = 7, walks: int = 16, noise=None):
        off = _u32(offsets)
        nb = _u32(neighbors)
        self._keep = (off, nb)
        nz = _need_len(None if noise is None else _f32(noise), 3 * (len(off) - 1), "noise")
        self.n = len(off) - 1
        _check(lib().jc_mpnn_set_graph(self._h, self.n, _ptr(off), _ptr(nb) if len(nb) else None,
                                       len(nb), feat_seed, walks, _ptr(nz)))

    def features(self):
        out = np.empty(self.n * 16, np.float32)
        _check(lib().jc_mpnn_features(self._h, _ptr(out)))
        return out.reshape(self.n, 16)

    def activations(self, layer: int):
        out = np.empty(self.n * 128, np.float32)
        _check(lib().jc_mpnn_activations(self._h, layer, _ptr(out)))
        return out.reshape(self.n, 128)

    def forward(self):
        """Returns (coordinates 3n, device ms)."""
        out = np.empty(3 * self.n, np.float32)
        ms = np.zeros(1, np.float64)
        _check(lib().jc_mpnn_forward(self._h, _ptr(out), _ptr(ms)))
        return out, float(ms[0])

    def forward_scaled(self, target_rms: float):
        """Coordinates (3n) centred at the origin and scaled to RMS radius
        target_rms on the device, and the device ms of forward plus rescale."""
        out = np.empty(3 * self.n, np.float32)
        ms = np.zeros(1, np.float64)
        _check(lib().jc_mpnn_forward_scaled(self._h, float(target_rms), _ptr(out), _ptr(ms)))
        return out, float(ms[0])

    def set_target(self, xyz, n_anchor: int = 32, seed: int = 3):
        x = _need_len(_f32(xyz), 3 * self.n, "target positions")
        _check(lib().jc_mpnn_set_target(self._h, _ptr(x), n_anchor, seed))

    def loss_grad(self) -> float:
        L = np.zeros(1, np.float64)
        _check(lib().jc_mpnn_loss_grad(self._h, _ptr(L)))
        return float(L[0])

    def adam(self, lr: float):
        _check(lib().jc_mpnn_adam(self._h, lr))

    def train(self, steps: int, lr: float) -> float:
        L = np.zeros(1, np.float64)
        _check(lib().jc_mpnn_train(self._h, steps, lr, _ptr(L)))
        return float(L[0])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().jc_mpnn_destroy(h)
            self._h = None


# ------------------------------------------------------------------ analytics
def pagerank(offsets, neighbors, damping: float = 0.85, max_iters: int = 100, tol_l1: float = 1e-6):
    """gpu_pagerank (crates/graph-compute/src/analytics/pagerank.rs:67-287) with an
    L1 stopping rule.  Returns (ranks, iterations, final L1 residual)."""
    offsets = _u32(offsets)
    neighbors = _u32(neighbors)
    n = len(offsets) - 1
    ranks = np.empty(max(n, 1), np.float32)
    it = np.zeros(1, np.uint32)
    res = np.zeros(1, np.float64)
    _check(lib().jc_pagerank(n, _ptr(offsets), _ptr(neighbors) if len(neighbors) else None,
                             len(neighbors), damping, max_iters, tol_l1, _ptr(ranks), _ptr(it),
                             _ptr(res)))
    return ranks[:n], int(it[0]), float(res[0])


PAGERANK_NOALLOC_MAX_N = 1 << 20


def pagerank_noalloc(n: int, d_offsets: int, d_neighbors: int, nnz: int, d_out_ranks: int,
                     damping: float = 0.85, max_iters: int = 100, tol_l1: float = 1e-6, stream: int = 0):
    """jc_pagerank_noalloc: PageRank over a CSR already in device memory (device
    addresses as integers), ranks written to d_out_ranks, no device allocation.
    Returns (iterations, final L1 residual)."""
    it = np.zeros(1, np.uint32)
    res = np.zeros(1, np.float64)
    _check(lib().jc_pagerank_noalloc(n, d_offsets or None, d_neighbors or None, nnz, damping, max_iters, tol_l1,
                                     d_out_ranks or None, _ptr(it), _ptr(res), stream or None))
    return int(it[0]), float(res[0])


class PageRankPlan:
    """A CSR graph resident on the device for repeated PageRank runs."""

    def __init__(self, offsets, neighbors, damping: float = 0.85):
        self._off = _u32(offsets)
        self._nb = _u32(neighbors)
        self.n = len(self._off) - 1
        h = C.c_void_p()
        _check(lib().jc_pagerank_plan_create(self.n, _ptr(self._off),
                                             _ptr(self._nb) if len(self._nb) else None,
                                             len(self._nb), damping, C.addressof(h)))
        self._h = h.value

    def run(self, max_iters: int = 100, tol_l1: float = 1e-6, readback: bool = True):
        """Returns (ranks or None, iterations, residual, device ms)."""
        ranks = np.empty(max(self.n, 1), np.float32) if readback else None
        it = np.zeros(1, np.uint32)
        res = np.zeros(1, np.float64)
        ms = np.zeros(1, np.float64)
        _check(lib().jc_pagerank_plan_run(self._h, max_iters, tol_l1, _ptr(ranks), _ptr(it),
                                          _ptr(res), _ptr(ms)))
        return (ranks[: self.n] if readback else None), int(it[0]), float(res[0]), float(ms[0])

    def time_iterations(self, n_iters: int) -> float:
        ms = np.zeros(1, np.float64)
        _check(lib().jc_pagerank_plan_time(self._h, n_iters, _ptr(ms)))
        return float(ms[0])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().jc_pagerank_plan_destroy(h)
            self._h = None


# ------------------------------------------------------------------ SpMV
SPMV_F32, SPMV_F16, SPMV_HYBRID = 0, 1, 2


def _spmv(fn, offsets, neighbors, weights, x):
    offsets, neighbors, weights, x = _u32(offsets), _u32(neighbors), _f32(weights), _f32(x)
    n = len(offsets) - 1
    if len(weights) != len(neighbors):
        raise JcError(JC_E_INVALID, f"neighbors ({len(neighbors)}) and weights ({len(weights)}) "
                                    "length mismatch")
    if len(x) != max(n, 0):
        raise JcError(JC_E_INVALID, f"x length {len(x)} != n ({n})")
    y = np.empty(max(n, 1), np.float32)
    _check(fn(n, _ptr(offsets), _ptr(neighbors) if len(neighbors) else None,
              _ptr(weights) if len(weights) else None, len(neighbors),
              _ptr(x) if n else None, _ptr(y)))
    return y[:n]


def spmv(offsets, neighbors, weights, x):
    """gpu_spmv (crates/graph-compute/src/analytics/spmv.rs:103-287): y = A x,
    each row summed in CSR order in f32."""
    return _spmv(lib().jc_spmv, offsets, neighbors, weights, x)


def spmv_f16(offsets, neighbors, weights, x):
    """gpu_spmv_f16 (spmv.rs:293-469): weights and x stored as f16 (nearest even),
    products and sums in f32."""
    return _spmv(lib().jc_spmv_f16, offsets, neighbors, weights, x)


def spmv_hybrid(offsets, neighbors, weights, x):
    """gpu_spmv_hybrid (spmv.rs:491-676): one 64-lane group per row, lane-strided
    partial sums and a stride-halving tree."""
    return _spmv(lib().jc_spmv_hybrid, offsets, neighbors, weights, x)


class SpmvPlan:
    """A weighted CSR matrix resident on the device for repeated products."""

    def __init__(self, offsets, neighbors, weights, kind: int = SPMV_F32):
        self._off, self._nb, self._w = _u32(offsets), _u32(neighbors), _f32(weights)
        self.n = len(self._off) - 1
        if len(self._w) != len(self._nb):
            raise JcError(JC_E_INVALID, "neighbors and weights length mismatch")
        h = C.c_void_p()
        _check(lib().jc_spmv_plan_create(self.n, _ptr(self._off),
                                         _ptr(self._nb) if len(self._nb) else None,
                                         _ptr(self._w) if len(self._w) else None,
                                         len(self._nb), kind, C.addressof(h)))
        self._h = h.value

    def run(self, x):
        """Returns (y, device ms of the product)."""
        x = _f32(x)
        if len(x) != self.n:
            raise JcError(JC_E_INVALID, f"x length {len(x)} != n ({self.n})")
        y = np.empty(max(self.n, 1), np.float32)
        ms = np.zeros(1, np.float64)
        _check(lib().jc_spmv_plan_run(self._h, _ptr(x) if self.n else None, _ptr(y), _ptr(ms)))
        return y[: self.n], float(ms[0])

    def time_products(self, n_iters: int) -> float:
        ms = np.zeros(1, np.float64)
        _check(lib().jc_spmv_plan_time(self._h, n_iters, _ptr(ms)))
        return float(ms[0])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().jc_spmv_plan_destroy(h)
            self._h = None


# ------------------------------------------------------------------ traversal
UNREACHABLE = 0xFFFFFFFF


def bfs(offsets, neighbors, source: int):
    """gpu_bfs (crates/graph-compute/src/analytics/bfs.rs:71-300): u32 hop
    distances from `source`, UNREACHABLE for nodes it does not reach.  Raises
    JcError (JC_E_INVALID) when source >= n; an empty graph gives an empty array."""
    offsets = _u32(offsets)
    neighbors = _u32(neighbors)
    n = len(offsets) - 1
    dist = np.empty(max(n, 1), np.uint32)
    _check(lib().jc_bfs(n, _ptr(offsets), _ptr(neighbors) if len(neighbors) else None,
                        len(neighbors), int(source), _ptr(dist)))
    return dist[:n]


class BfsPlan:
    """A CSR graph resident on the device for repeated single-source BFS."""

    def __init__(self, offsets, neighbors):
        self._off = _u32(offsets)
        self._nb = _u32(neighbors)
        self.n = len(self._off) - 1
        h = C.c_void_p()
        _check(lib().jc_bfs_plan_create(self.n, _ptr(self._off), _ptr(self._nb) if len(self._nb) else None,
                                        len(self._nb), C.addressof(h)))
        self._h = h.value

    def run(self, source: int, readback: bool = True):
        """Returns (distances or None, device ms of the search)."""
        dist = np.empty(max(self.n, 1), np.uint32) if readback else None
        ms = np.zeros(1, np.float64)
        _check(lib().jc_bfs_plan_run(self._h, int(source), _ptr(dist), _ptr(ms)))
        return (dist[: self.n] if readback else None), float(ms[0])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().jc_bfs_plan_destroy(h)
            self._h = None


def connected_components(offsets, neighbors):
    """gpu_connected_components (crates/graph-compute/src/analytics/
    connected_components.rs:74): label = smallest node index of the node's
    component, CSR entries taken as undirected edges."""
    offsets = _u32(offsets)
    neighbors = _u32(neighbors)
    n = len(offsets) - 1
    lab = np.empty(max(n, 1), np.uint32)
    _check(lib().jc_connected_components(n, _ptr(offsets), _ptr(neighbors) if len(neighbors) else None,
                                         len(neighbors), _ptr(lab)))
    return lab[:n]


def largest_component(offsets, neighbors):
    """The largest component (ties: the one holding the smallest node id) on the
    device.  Returns (keep, sub_offsets, sub_neighbors): keep[v] = new id of v,
    increasing in v, or UNREACHABLE for nodes outside it; the sub-CSR is the
    induced graph with renumbered ids."""
    offsets = _u32(offsets)
    neighbors = _u32(neighbors)
    n = len(offsets) - 1
    keep = np.empty(max(n, 1), np.uint32)
    size = np.zeros(1, np.uint32)
    sub_off = np.empty(n + 1, np.uint32)
    sub_nb = np.empty(max(len(neighbors), 1), np.uint32)
    sub_nnz = np.zeros(1, np.uint64)
    _check(lib().jc_largest_component(n, _ptr(offsets), _ptr(neighbors) if len(neighbors) else None,
                                      len(neighbors), _ptr(keep), _ptr(size), _ptr(sub_off), _ptr(sub_nb),
                                      _ptr(sub_nnz)))
    k = int(size[0])
    return keep[:n], sub_off[: k + 1].copy(), sub_nb[: int(sub_nnz[0])].copy()


# ------------------------------------------------------------------ fixtures
def csr_from_pairs(n: int, pairs):
    pairs = _u32(pairs).reshape(-1)
    offsets = np.empty(n + 1, np.uint32)
    nbrs = np.empty(max(len(pairs), 1), np.uint32)
    nnz = lib().jc_csr_from_pairs(n, _ptr(pairs), len(pairs) // 2, _ptr(offsets), _ptr(nbrs))
    return offsets, nbrs[:nnz].copy()


def ba_pairs(n: int, m: int, seed: int):
    cap = 1 + (max(n, 2) - 2) * max(m, 1)
    pairs = np.empty(2 * cap, np.uint32)
    cnt = lib().jc_gen_ba_pairs(n, m, seed, _ptr(pairs), cap)
    return pairs[: 2 * cnt]


def ba_graph(n: int, m: int, seed: int):
    """synthetic_scale_free (crates/graph-compute/src/stability.rs:57-82) as CSR."""
    return csr_from_pairs(max(n, 2), ba_pairs(n, m, seed))


def sbm_pairs(n, blocks, p_in, p_out, seed):
    cap = 1 << 16
    while True:
        pairs = np.empty(2 * cap, np.uint32)
        cnt = lib().jc_gen_sbm_pairs(n, blocks, p_in, p_out, seed, _ptr(pairs), cap)
        if cnt <= cap:
            return pairs[: 2 * cnt]
        cap = int(cnt)


def sbm_graph(n, blocks, p_in, p_out, seed):
    return csr_from_pairs(n, sbm_pairs(n, blocks, p_in, p_out, seed))


def rgg_pairs(n, radius, seed):
    xy = np.empty(2 * max(n, 1), np.float32)
    cap = max(16, n * 6)
    while True:
        pairs = np.empty(2 * cap, np.uint32)
        cnt = lib().jc_gen_rgg_pairs(n, radius, seed, _ptr(xy), _ptr(pairs), cap)
        if cnt <= cap:
            return xy[: 2 * n], pairs[: 2 * cnt]
        cap = int(cnt)


def _splitmix_pair(seed: int):
    """First two SplitMix64 outputs of `seed`: (u in [0,1) from the top 53 bits, h)."""
    M = (1 << 64) - 1
    st, outs = seed & M, []
    for _ in range(2):
        st = (st + 0x9E3779B97F4A7C15) & M
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        outs.append(z ^ (z >> 31))
    return (outs[0] >> 11) / float(1 << 53), outs[1]


def uniform_cube(n: int, half: float, seed: int):
    out = np.empty(3 * max(n, 1), np.float32)
    lib().jc_gen_uniform_cube(n, half, seed, _ptr(out))
    return out[: 3 * n]


def ball_seed(n: int, seed: int):
    """ball_seed (stability.rs:30-50): uniform random ball of radius sqrt(n)*5."""
    out = np.empty(3 * max(n, 1), np.float32)
    lib().jc_gen_ball_seed(n, seed, _ptr(out))
    return out[: 3 * n]


class StepStats(C.Structure):
    """StepStats (stability.rs:128-186)."""

    _fields_ = [("max_abs", C.c_float), ("p50_radius", C.c_float), ("p99_radius", C.c_float),
                ("mean_disp", C.c_float), ("p50_disp", C.c_float), ("nonfinite", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def step_stats(prev, cur) -> StepStats:
    """StepStats::measure(prev, cur): per-step stability telemetry."""
    prev, cur = _f32(prev), _f32(cur)
    if len(prev) != len(cur) or len(cur) % 3:
        raise JcError(JC_E_INVALID, f"positions lengths {len(prev)} / {len(cur)} do not match 3n")
    st = StepStats()
    _check(lib().jc_step_stats(len(cur) // 3, _ptr(prev) if len(prev) else None,
                               _ptr(cur) if len(cur) else None, C.addressof(st)))
    return st


def csr_from_bin(data) -> tuple:
    """CsrGraph::from_bin_bytes (sim.rs:72-104): (offsets, neighbors) from the
    little-endian binary CSR; raises JcError on a truncated or mis-sized buffer."""
    buf = np.frombuffer(bytes(data), np.uint8)
    n = np.zeros(1, np.uint32)
    m = np.zeros(1, np.uint32)
    _check(lib().jc_csr_bin_header(_ptr(buf) if len(buf) else None, len(buf), _ptr(n), _ptr(m)))
    off = np.empty(int(n[0]) + 1, np.uint32)
    nb = np.empty(max(int(m[0]), 1), np.uint32)
    _check(lib().jc_csr_from_bin(_ptr(buf), len(buf), _ptr(off), _ptr(nb)))
    return off, nb[: int(m[0])]


def csr_to_bin(offsets, neighbors) -> bytes:
    """CsrGraph::to_bin (sim.rs:118-127)."""
    off = _u32(offsets)
    nb = _u32(neighbors)
    n = len(off) - 1
    size = int(lib().jc_csr_bin_size(n, len(nb)))
    out = np.empty(size, np.uint8)
    _check(lib().jc_csr_to_bin(n, _ptr(off), _ptr(nb) if len(nb) else None, len(nb), _ptr(out), size))
    return out.tobytes()


def load_csr_bin(path) -> tuple:
    """CsrGraph::load_bin (sim.rs:61-66)."""
    with open(path, "rb") as fh:
        return csr_from_bin(fh.read())


def write_csr_bin(path, offsets, neighbors) -> None:
    """CsrGraph::write_bin (sim.rs:107-113)."""
    with open(path, "wb") as fh:
        fh.write(csr_to_bin(offsets, neighbors))
