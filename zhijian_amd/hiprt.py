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
