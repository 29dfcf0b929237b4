"""Device buffers for the tests of entry points that take device pointers:
hipMalloc / hipMemcpy / hipFree through ctypes on the HIP runtime that
libjc_hip.so itself links (no torch)."""
import ctypes as C

import numpy as np

_H2D, _D2H = 1, 2
_rt = None


def runtime():
    global _rt
    if _rt is None:
        rt = C.CDLL("libamdhip64.so.7")
        rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        rt.hipFree.argtypes = [C.c_void_p]
        rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        rt.hipDeviceSynchronize.argtypes = []
        for f in (rt.hipMalloc, rt.hipFree, rt.hipMemcpy, rt.hipDeviceSynchronize):
            f.restype = C.c_int
        _rt = rt
    return _rt


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """A device copy of a numpy array (or an uninitialised one of a given size)."""

    def __init__(self, host=None, *, dtype=None, count=None):
        if host is not None:
            host = np.ascontiguousarray(host)
            dtype, count = host.dtype, host.size
        self.dtype, self.count = np.dtype(dtype), int(count)
        self.nbytes = self.dtype.itemsize * self.count
        p = C.c_void_p()
        _ok(runtime().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = int(p.value)
        if host is not None and self.nbytes:
            _ok(runtime().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, _H2D), "hipMemcpy H2D")

    def numpy(self):
        out = np.empty(self.count, self.dtype)
        _ok(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")
        if self.nbytes:
            _ok(runtime().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
