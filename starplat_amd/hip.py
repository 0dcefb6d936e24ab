"""Minimal ctypes access to the HIP runtime (libamdhip64): device buffers,
copies and events. This is plumbing for callers that keep results in HBM
(benchmarks, batched jobs); the kernels themselves live behind the C ABI.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_lib = None
H2D, D2H, D2D = 1, 2, 3


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        _lib = C.CDLL(path if os.path.exists(path) else "libamdhip64.so", mode=C.RTLD_GLOBAL)
        _lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _lib.hipFree.argtypes = [C.c_void_p]
        _lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _lib.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        _lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        _lib.hipDeviceSynchronize.argtypes = []
        _lib.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        _lib.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _lib.hipEventDestroy.argtypes = [C.c_void_p]
        _lib.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _lib.hipEventSynchronize.argtypes = [C.c_void_p]
        _lib.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _lib.hipGetErrorString.restype = C.c_char_p
        _lib.hipGetErrorString.argtypes = [C.c_int]
    return _lib


def _chk(name, rc):
    if rc != 0:
        raise RuntimeError(f"{name}: {lib().hipGetErrorString(rc).decode()}")


def device_count():
    n = C.c_int(0)
    rc = lib().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    _chk("hipDeviceSynchronize", lib().hipDeviceSynchronize())


class DeviceArray:
    """A device allocation of n elements of a numpy dtype."""

    def __init__(self, n, dtype):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = max(self.n * self.dtype.itemsize, 4)
        p = C.c_void_p()
        _chk("hipMalloc", lib().hipMalloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    def to_host(self):
        out = np.empty(self.n, self.dtype)
        if self.n:
            _chk("hipMemcpy", lib().hipMemcpy(out.ctypes.data, self.ptr, self.n * self.dtype.itemsize, D2H))
        return out

    def from_host(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.n
        if self.n:
            _chk("hipMemcpy", lib().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D))
        return self

    def free(self):
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        p = C.c_void_p()
        _chk("hipEventCreate", lib().hipEventCreate(C.byref(p)))
        self.ev = p.value

    def record(self, stream):
        _chk("hipEventRecord", lib().hipEventRecord(self.ev, stream))

    def synchronize(self):
        _chk("hipEventSynchronize", lib().hipEventSynchronize(self.ev))

    def elapsed_ms(self, end):
        ms = C.c_float(0)
        _chk("hipEventElapsedTime", lib().hipEventElapsedTime(C.byref(ms), self.ev, end.ev))
        return ms.value

    def __del__(self):
        try:
            lib().hipEventDestroy(self.ev)
        except Exception:
            pass
