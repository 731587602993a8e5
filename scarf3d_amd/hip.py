"""Minimal device-memory plumbing for tests and benchmarks.

Binds the HIP runtime the library itself links (libamdhip64.so.7 from the
ROCm installation) with ctypes: allocation, copies, synchronisation and
events. A second HIP runtime in the same process (for example the one
bundled inside a PyTorch wheel) cannot open the GPU once the first has, so
nothing here goes through PyTorch.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_H2D, _D2H, _D2D = 1, 2, 3


def _load():
    from . import api

    # loading the library first makes the dynamic loader resolve its runtime
    api.load("f")
    for name in ("libamdhip64.so.7", os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so.7")):
        try:
            return C.CDLL(name)
        except OSError:
            continue
    raise OSError("libamdhip64.so.7 not found")


class Hip:
    def __init__(self):
        L = _load()
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.hipDeviceSynchronize.argtypes = []
        L.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        L.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.hipGetErrorString.restype = C.c_char_p
        self.L = L

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: {self.L.hipGetErrorString(rc).decode()}")

    def device_count(self) -> int:
        n = C.c_int(0)
        rc = self.L.hipGetDeviceCount(C.byref(n))
        return n.value if rc == 0 else 0

    def mem_get_info(self):
        """(free, total) bytes of the current device (hipMemGetInfo)."""
        free, total = C.c_size_t(0), C.c_size_t(0)
        self._check(self.L.hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
        return free.value, total.value

    def synchronize(self):
        self._check(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def empty(self, n, dtype):
        return DeviceArray(self, int(n), np.dtype(dtype))


class DeviceArray:
    """A device buffer of n elements of dtype."""

    def __init__(self, hip: Hip, n: int, dtype: np.dtype):
        self.hip, self.n, self.dtype = hip, n, dtype
        self.nbytes = n * dtype.itemsize
        p = C.c_void_p()
        hip._check(hip.L.hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        self.hip._check(self.hip.L.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype).reshape(-1)
        assert a.size == self.n
        self.hip._check(self.hip.L.hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, _H2D), "hipMemcpy H2D")

    def zero(self):
        self.hip._check(self.hip.L.hipMemset(self.ptr, 0, self.nbytes), "hipMemset")
        self.hip.synchronize()

    def free(self):
        if self.ptr:
            self.hip.L.hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        # at interpreter exit the runtime may already be tearing down: leave
        # the memory to the process exit instead of calling into HIP then
        if sys.is_finalizing():
            return
        try:
            self.free()
        except Exception:
            pass
