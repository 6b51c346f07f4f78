"""Minimal HIP runtime access (device buffers, copies, synchronisation) through
libamdhip64, so benchmarks can keep inputs and outputs resident in HBM without
importing a framework."""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        h.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        _hip = h
    return _hip


H2D, D2H, D2D = 1, 2, 3


def check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class DeviceBuffer:
    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        check(hip().hipMalloc(C.byref(self.ptr), self.nbytes), "hipMalloc")

    @property
    def addr(self) -> int:
        return self.ptr.value

    def to_host(self, dtype=np.float32, shape=None) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out.reshape(shape) if shape is not None else out

    def from_host(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def device_count() -> int:
    n = C.c_int(0)
    if hip().hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value
