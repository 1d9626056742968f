"""Minimal ctypes access to the HIP runtime's streams and events, for timing
launches on the stream they run on (no torch import needed)."""
from __future__ import annotations

import ctypes as C
import os

_hip = None


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        for cand in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(cand)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        H = _hip
        H.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        H.hipStreamDestroy.argtypes = [C.c_void_p]
        H.hipStreamSynchronize.argtypes = [C.c_void_p]
        H.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        H.hipEventDestroy.argtypes = [C.c_void_p]
        H.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        H.hipEventSynchronize.argtypes = [C.c_void_p]
        H.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        H.hipDeviceSynchronize.argtypes = []
        H.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class Stream:
    def __init__(self):
        self.h = C.c_void_p()
        _ok(hip().hipStreamCreate(C.byref(self.h)), "hipStreamCreate")

    @property
    def ptr(self):
        return self.h.value

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.h:
                hip().hipStreamDestroy(self.h)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        _ok(hip().hipEventCreate(C.byref(self.h)), "hipEventCreate")

    def record(self, stream: Stream):
        _ok(hip().hipEventRecord(self.h, stream.h), "hipEventRecord")

    def synchronize(self):
        _ok(hip().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        v = C.c_float()
        _ok(hip().hipEventElapsedTime(C.byref(v), self.h, end.h), "hipEventElapsedTime")
        return float(v.value)

    def __del__(self):
        try:
            if self.h:
                hip().hipEventDestroy(self.h)
        except Exception:
            pass


def device_synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def device_count() -> int:
    n = C.c_int(0)
    rc = hip().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


_H2D, _D2H = 1, 2


class DeviceArray:
    """A device buffer of `n` items of a numpy dtype (hipMalloc / hipMemcpy),
    filled from a host array or zeroed."""

    def __init__(self, n_or_host, dtype=None):
        import numpy as np
        if isinstance(n_or_host, (int, np.integer)):
            self.dtype = np.dtype(dtype or np.float64)
            self.n = int(n_or_host)
            host = None
        else:
            host = np.ascontiguousarray(n_or_host, dtype=dtype or n_or_host.dtype)
            self.dtype = host.dtype
            self.n = host.size
        self.nbytes = max(self.n * self.dtype.itemsize, 8)
        self.h = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(self.h), self.nbytes), "hipMalloc")
        if host is not None and host.size:
            _ok(hip().hipMemcpy(self.h, host.ctypes.data, host.nbytes, _H2D), "hipMemcpy")
        elif host is None:
            _ok(hip().hipMemset(self.h, 0, self.nbytes), "hipMemset")

    @property
    def ptr(self):
        return self.h.value

    def upload(self, host):
        import numpy as np
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.size == self.n
        _ok(hip().hipMemcpy(self.h, host.ctypes.data, host.nbytes, _H2D), "hipMemcpy")

    def get(self):
        import numpy as np
        out = np.empty(self.n, self.dtype)
        if self.n:
            _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
            _ok(hip().hipMemcpy(out.ctypes.data, self.h, out.nbytes, _D2H), "hipMemcpy")
        return out

    def __del__(self):
        try:
            if self.h:
                hip().hipFree(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass
