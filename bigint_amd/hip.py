"""Minimal ctypes access to the HIP runtime that libbigint_hip.so itself links
(/opt/rocm/lib/libamdhip64.so.7): device buffers, copies, streams and events for the tests,
bench.py and the planner.  Using the same runtime as the engine keeps one HIP runtime per
process (PyTorch wheels bundle a second one)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PATHS = ["/opt/rocm/lib/libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"]
_rt = None

H2D, D2H, D2D = 1, 2, 3


def rt() -> ctypes.CDLL:
    global _rt
    if _rt is None:
        for p in _PATHS:
            if os.path.exists(p):
                _rt = ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)
                break
        else:
            raise FileNotFoundError("HIP runtime not found under /opt/rocm/lib")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        _rt.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        _rt.hipFree.argtypes = [vp]
        _rt.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        _rt.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        _rt.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        _rt.hipStreamSynchronize.argtypes = [vp]
        _rt.hipStreamDestroy.argtypes = [vp]
        _rt.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        _rt.hipEventRecord.argtypes = [vp, vp]
        _rt.hipEventSynchronize.argtypes = [vp]
        _rt.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        _rt.hipEventDestroy.argtypes = [vp]
        _rt.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
    return _rt


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


def device_count() -> int:
    n = ctypes.c_int(0)
    if rt().hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


def synchronize() -> None:
    _check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Buffer:
    """Device allocation of `nbytes` bytes; .ptr is the raw device address."""

    def __init__(self, nbytes: int):
        p = ctypes.c_void_p()
        _check(rt().hipMalloc(ctypes.byref(p), max(1, nbytes)), "hipMalloc")
        self.ptr = p.value
        self.nbytes = nbytes

    def free(self) -> None:
        if self.ptr:
            rt().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @classmethod
    def from_array(cls, a: np.ndarray) -> "Buffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        _check(rt().hipMemcpy(ctypes.c_void_p(self.ptr + offset), a.ctypes.data_as(ctypes.c_void_p),
                              a.nbytes, H2D), "hipMemcpy H2D")

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        _check(rt().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr + offset),
                              out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self) -> None:
        _check(rt().hipMemset(ctypes.c_void_p(self.ptr), 0, self.nbytes), "hipMemset")


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        _check(rt().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self) -> None:
        _check(rt().hipStreamSynchronize(ctypes.c_void_p(self.handle)), "hipStreamSynchronize")


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _check(rt().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream: Stream) -> None:
        _check(rt().hipEventRecord(ctypes.c_void_p(self.handle), ctypes.c_void_p(stream.handle)),
               "hipEventRecord")

    def synchronize(self) -> None:
        _check(rt().hipEventSynchronize(ctypes.c_void_p(self.handle)), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        _check(rt().hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(self.handle),
                                        ctypes.c_void_p(end.handle)), "hipEventElapsedTime")
        return float(ms.value)
