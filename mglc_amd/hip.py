"""Minimal ctypes access to the HIP runtime that libmglc_lbm.so itself links
(libamdhip64.so.7): streams and events for timing.  Loading the product
library first makes the dynamic loader hand back that same runtime, so
events and streams here are valid for the library's launches."""
from __future__ import annotations

import ctypes

from . import lbm

_hip = None


def runtime() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        lbm.load_library()
        h = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        h.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipEventRecord.argtypes = [vp, vp]
        h.hipEventSynchronize.argtypes = [vp]
        h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        h.hipEventDestroy.argtypes = [vp]
        h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipStreamDestroy.argtypes = [vp]
        h.hipStreamSynchronize.argtypes = [vp]
        h.hipGetErrorString.restype = ctypes.c_char_p
        _hip = h
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {runtime().hipGetErrorString(rc).decode()}")


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        _ok(runtime().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    def synchronize(self):
        _ok(runtime().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self):
        if self.handle:
            runtime().hipStreamDestroy(self.handle)
            self.handle = ctypes.c_void_p()


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        _ok(runtime().hipEventCreate(ctypes.byref(self.handle)), "hipEventCreate")

    def record(self, stream: Stream | None):
        _ok(runtime().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self):
        _ok(runtime().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        _ok(runtime().hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return float(ms.value)

    def close(self):
        if self.handle:
            runtime().hipEventDestroy(self.handle)
            self.handle = ctypes.c_void_p()
