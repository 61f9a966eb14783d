"""Minimal HIP runtime binding for device buffers, streams and events.

Device memory for the C-ABI is plain `hipMalloc` memory owned by the caller;
this module is the Python caller's plumbing (no framework tensors involved),
loaded from the same libamdhip64 that libpvx.so links against.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

_H = None

hipMemcpyHostToDevice = 1
hipMemcpyDeviceToHost = 2
hipMemcpyDeviceToDevice = 3


def _lib():
    global _H
    if _H is None:
        for name in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so.7", "libamdhip64.so"):
            try:
                _H = C.CDLL(name, mode=C.RTLD_GLOBAL)
                break
            except OSError:
                continue
        if _H is None:
            raise RuntimeError("HIP runtime (libamdhip64) not found")
        _H.hipGetErrorString.restype = C.c_char_p
        _H.hipGetErrorString.argtypes = [C.c_int]
    return _H


def check(rc, what="hip call"):
    if rc != 0:
        msg = _lib().hipGetErrorString(rc).decode()
        raise RuntimeError(f"{what} failed: {msg} ({rc})")


def device_count() -> int:
    n = C.c_int(0)
    rc = _lib().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    check(_lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """A hipMalloc'd buffer with numpy upload/download helpers."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        if self.nbytes > 0:
            check(_lib().hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)), "hipMalloc")

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    @property
    def value(self):
        return self.ptr.value

    def upload(self, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            check(_lib().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes),
                                   hipMemcpyHostToDevice), "hipMemcpy H2D")

    def download(self, dtype, count=None) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype)
        if out.nbytes:
            check(_lib().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr,
                                   C.c_size_t(out.nbytes), hipMemcpyDeviceToHost),
                  "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr and self.ptr.value:
            _lib().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.h = C.c_void_p()
        check(_lib().hipStreamCreate(C.byref(self.h)), "hipStreamCreate")

    def synchronize(self):
        check(_lib().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def destroy(self):
        if self.h and self.h.value:
            _lib().hipStreamDestroy(self.h)
            self.h = C.c_void_p()


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        check(_lib().hipEventCreate(C.byref(self.h)), "hipEventCreate")

    def record(self, stream: Stream | None = None):
        check(_lib().hipEventRecord(self.h, stream.h if stream else None), "hipEventRecord")

    def synchronize(self):
        check(_lib().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(_lib().hipEventElapsedTime(C.byref(ms), self.h, end.h), "hipEventElapsedTime")
        return ms.value

    def destroy(self):
        if self.h and self.h.value:
            _lib().hipEventDestroy(self.h)
            self.h = C.c_void_p()
