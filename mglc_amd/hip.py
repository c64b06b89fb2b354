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
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        h.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
        _hip = h
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {runtime().hipGetErrorString(rc).decode()}")


class Stream:
    def __init__(self, borrowed: int | None = None):
        """A new stream, or (borrowed = a hipStream_t address) a reference to a
        stream owned elsewhere, which close() leaves alone."""
        self.handle = ctypes.c_void_p()
        self._owned = borrowed is None
        if self._owned:
            _ok(runtime().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")
        else:
            self.handle = ctypes.c_void_p(borrowed)

    def synchronize(self):
        _ok(runtime().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self):
        if self.handle and self._owned:
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


def device_copy_calibration(nbytes: int, reps: int) -> None:
    """reps device-to-device copies of nbytes (HIP's blit kernel): a plain copy
    of known size, against which the HBM byte counters are scaled."""
    h = runtime()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    _ok(h.hipMalloc(ctypes.byref(a), nbytes), "hipMalloc")
    _ok(h.hipMalloc(ctypes.byref(b), nbytes), "hipMalloc")
    _ok(h.hipMemset(a, 1, nbytes), "hipMemset")
    for _ in range(reps):
        _ok(h.hipMemcpy(b, a, nbytes, 3), "hipMemcpy")  # 3 = hipMemcpyDeviceToDevice
    h.hipFree(a)
    h.hipFree(b)


class DeviceBuffer:
    """A plain device allocation (for halo send/receive buffers in tests and
    drivers that exchange between subdomains)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        _ok(runtime().hipMalloc(ctypes.byref(self.ptr), max(1, self.nbytes)), "hipMalloc")

    @property
    def address(self) -> int:
        return int(self.ptr.value)

    def copy_from(self, src: "DeviceBuffer", dst_offset: int, src_offset: int, nbytes: int):
        if dst_offset + nbytes > self.nbytes or src_offset + nbytes > src.nbytes or min(dst_offset, src_offset) < 0:
            raise ValueError("device copy out of bounds")
        _ok(runtime().hipMemcpy(ctypes.c_void_p(self.address + dst_offset), ctypes.c_void_p(src.address + src_offset),
                                nbytes, 3), "hipMemcpy")

    def close(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()
