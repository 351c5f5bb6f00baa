"""ctypes binding of the C ABI in include/bigint_hip.h (libbigint_hip.so, built in-tree).

The shared library must exist; nothing here falls back to a CPU path.  Tests and bench.py call
the engine through this module, i.e. through the same C ABI a C/C++ caller uses.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbigint_hip.so")

OK, EINVAL, ETOOBIG, ENOMEM, EDEVICE, ENODEV, ECAP, ESCRATCH, ENOINIT = range(9)
PLAN_FIXED, PLAN_TABLE, PLAN_LEARNED = 0, 1, 2

EXPORTS = ["bi_hip_init", "bi_hip_mul", "bi_hip_mul_top", "bi_hip_mul_device", "bi_hip_mul_batch",
           "bi_hip_synchronize", "bi_hip_transform_log2", "bi_hip_strerror", "bi_hip_profile",
           "bi_hip_profile_kinds",
           "bi_hip_profile_read", "bi_hip_mul_batch_device", "bi_hip_plan_count",
           "bi_hip_plan_allowed", "bi_hip_plan_cost", "bi_hip_mul_batch_device_planned", "bi_hip_choose_plans",
           "bi_hip_set_planner", "bi_hip_load_planner", "bi_hip_mlp_param_count",
           "bi_hip_mlp_forward", "bi_hip_mlp_grad", "bi_hip_mlp_adam", "bi_hip_set_graphs",
           "bi_hip_add_batch_device", "bi_hip_device_alloc", "bi_hip_device_free",
           "bi_hip_copy_to_device", "bi_hip_copy_to_host", "bi_hip_mul_device_noalloc"]
KERNEL_NAMES = ["fill", "outer_fwd", "mid", "outer_inv", "garner_carry", "outer_first"]

u32p = ctypes.POINTER(ctypes.c_uint32)


class DAdd(ctypes.Structure):
    _fields_ = [("x", ctypes.c_void_p), ("lx", ctypes.c_size_t), ("y", ctypes.c_void_p),
                ("ly", ctypes.c_size_t), ("subtract", ctypes.c_int), ("out", ctypes.c_void_p),
                ("out_cap", ctypes.c_size_t), ("swapped", ctypes.c_int), ("out_len", ctypes.c_size_t)]


class Pair(ctypes.Structure):
    _fields_ = [("a", u32p), ("la", ctypes.c_size_t), ("b", u32p), ("lb", ctypes.c_size_t),
                ("out", u32p), ("out_cap", ctypes.c_size_t), ("out_len", ctypes.c_size_t),
                ("status", ctypes.c_int)]


class DPair(ctypes.Structure):
    _fields_ = [("a", ctypes.c_void_p), ("la", ctypes.c_size_t), ("b", ctypes.c_void_p),
                ("lb", ctypes.c_size_t), ("out", ctypes.c_void_p), ("out_cap", ctypes.c_size_t)]


class KStat(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double), ("bytes", ctypes.c_double), ("launches", ctypes.c_uint64)]


class EngineError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        super().__init__(f"{what}: {strerror(code)} ({code})" if what else strerror(code))


_lib = None


def lib() -> ctypes.CDLL:
    """Load libbigint_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(LIB_PATH)
        L.bi_hip_init.argtypes = [ctypes.c_int]
        L.bi_hip_mul.argtypes = [u32p, ctypes.c_size_t, u32p, ctypes.c_size_t, ctypes.c_int, u32p,
                                 ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.bi_hip_mul_top.argtypes = [u32p, ctypes.c_size_t, u32p, ctypes.c_size_t, ctypes.c_int,
                                     ctypes.c_size_t, u32p, ctypes.POINTER(ctypes.c_size_t)]
        L.bi_hip_mul_device.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                        ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_size_t, ctypes.c_void_p]
        L.bi_hip_mul_device_noalloc.argtypes = L.bi_hip_mul_device.argtypes
        L.bi_hip_mul_batch.argtypes = [ctypes.POINTER(Pair), ctypes.c_size_t, ctypes.c_int]
        L.bi_hip_mul_batch_device.argtypes = [ctypes.POINTER(DPair), ctypes.c_size_t, ctypes.c_int,
                                              ctypes.c_void_p]
        L.bi_hip_transform_log2.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
        L.bi_hip_profile.argtypes = [ctypes.c_int]
        L.bi_hip_profile_kinds.argtypes = [ctypes.c_uint32]
        L.bi_hip_profile_read.argtypes = [ctypes.POINTER(KStat), ctypes.c_size_t]
        vp = ctypes.c_void_p
        L.bi_hip_plan_allowed.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
        L.bi_hip_mlp_param_count.argtypes = [ctypes.c_int]
        L.bi_hip_mlp_forward.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, vp, vp]
        L.bi_hip_mlp_grad.argtypes = [vp, vp, vp, ctypes.c_int, vp, ctypes.c_int, vp, vp]
        L.bi_hip_mlp_adam.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_int, vp]
        L.bi_hip_add_batch_device.argtypes = [ctypes.POINTER(DAdd), ctypes.c_size_t, ctypes.c_void_p]
        L.bi_hip_device_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]
        L.bi_hip_device_free.argtypes = [ctypes.c_void_p]
        L.bi_hip_copy_to_device.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                            ctypes.POINTER(ctypes.c_size_t), ctypes.c_size_t]
        L.bi_hip_copy_to_host.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t),
                                          ctypes.c_size_t, ctypes.c_void_p]
        L.bi_hip_strerror.restype = ctypes.c_char_p
        L.bi_hip_strerror.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def strerror(code: int) -> str:
    return lib().bi_hip_strerror(code).decode()


def init(device: int = 0) -> None:
    rc = lib().bi_hip_init(device)
    if rc:
        raise EngineError(rc, "bi_hip_init")


def transform_log2(la: int, lb: int) -> int:
    return int(lib().bi_hip_transform_log2(la, lb))


def _p(a: np.ndarray):
    return a.ctypes.data_as(u32p)


def mul(a: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    """Product of two 28-bit limb arrays (host memory); b=None squares a."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    sq = b is None
    bb = a if sq else np.ascontiguousarray(b, dtype=np.uint32)
    out = np.zeros(a.size + bb.size, dtype=np.uint32)
    n = ctypes.c_size_t()
    rc = lib().bi_hip_mul(_p(a), a.size, _p(bb), bb.size, int(sq), _p(out), out.size, ctypes.byref(n))
    if rc:
        raise EngineError(rc, "bi_hip_mul")
    return out[: n.value]


def mul_top(a: np.ndarray, b: np.ndarray | None, keep: int) -> tuple[np.ndarray, int]:
    """The top min(keep, |a| + |b|) limbs of a * b (b=None squares a) and the length of the whole
    product without leading zeros."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    sq = b is None
    bb = a if sq else np.ascontiguousarray(b, dtype=np.uint32)
    out = np.zeros(max(1, min(keep, a.size + bb.size)), dtype=np.uint32)
    n = ctypes.c_size_t()
    rc = lib().bi_hip_mul_top(_p(a), a.size, _p(bb), bb.size, int(sq), keep, _p(out), ctypes.byref(n))
    if rc:
        raise EngineError(rc, "bi_hip_mul_top")
    return out, n.value


def add_batch_device(jobs: list[tuple[int, int, int, int, bool, int, int]]) -> list[tuple[bool, int]]:
    """Device-resident signed-magnitude additions: jobs of (d_x, lx, d_y, ly, subtract, d_out,
    out_cap) with raw device addresses; returns (swapped, out_len) per job."""
    arr = (DAdd * len(jobs))()
    for i, (x, lx, y, ly, sub, out, cap) in enumerate(jobs):
        arr[i] = DAdd(x, lx, y, ly, int(sub), out, cap, 0, 0)
    rc = lib().bi_hip_add_batch_device(arr, len(jobs), None)
    if rc:
        raise EngineError(rc, "bi_hip_add_batch_device")
    return [(bool(a.swapped), int(a.out_len)) for a in arr]


def mul_device(d_a: int, la: int, d_b: int, lb: int, square: bool, d_out: int, out_cap: int,
               stream: int = 0) -> None:
    """Device-resident product; pointers are raw device addresses (e.g. tensor.data_ptr())."""
    rc = lib().bi_hip_mul_device(d_a, la, d_b, lb, int(square), d_out, out_cap, stream or None)
    if rc:
        raise EngineError(rc, "bi_hip_mul_device")


def mul_batch(pairs: list[tuple[np.ndarray, np.ndarray | None]], plan_mode: int = PLAN_FIXED):
    """Batch of products on host arrays; returns the list of products."""
    arr = (Pair * len(pairs))()
    keep, outs = [], []
    for i, (a, b) in enumerate(pairs):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        bb = a if b is None else np.ascontiguousarray(b, dtype=np.uint32)
        out = np.zeros(a.size + bb.size, dtype=np.uint32)
        keep += [a, bb]
        outs.append(out)
        arr[i] = Pair(_p(a), a.size, _p(bb), bb.size, _p(out), out.size, 0, 0)
    rc = lib().bi_hip_mul_batch(arr, len(pairs), plan_mode)
    if rc:
        raise EngineError(rc, "bi_hip_mul_batch")
    return [outs[i][: arr[i].out_len] for i in range(len(pairs))]


def mul_batch_device(pairs: list[tuple[int, int, int, int, int, int]], plan_mode: int = PLAN_TABLE,
                     stream: int = 0) -> None:
    """Device batch: each pair is (d_a, la, d_b or 0, lb, d_out, out_cap) with raw device pointers.
    The pairs go to the engine as one array of six 64-bit words each (bi_hip_dpair's layout: three
    pointers and three sizes, a null b meaning a square)."""
    try:
        arr = np.array(pairs, dtype=np.uint64).reshape(-1, 6)
    except TypeError:  # a None for d_b
        arr = np.array([(a, la, b or 0, lb, o, cap) for a, la, b, lb, o, cap in pairs],
                       dtype=np.uint64).reshape(-1, 6)
    assert ctypes.sizeof(DPair) == 48
    rc = lib().bi_hip_mul_batch_device(arr.ctypes.data_as(ctypes.POINTER(DPair)), len(arr), plan_mode,
                                       stream or None)
    if rc:
        raise EngineError(rc, "bi_hip_mul_batch_device")


def synchronize() -> None:
    rc = lib().bi_hip_synchronize()
    if rc:
        raise EngineError(rc, "bi_hip_synchronize")


def profile(enable: bool) -> None:
    lib().bi_hip_profile(int(enable))


def profile_kinds(names) -> None:
    """Record events only around launches of the named kernel classes (an empty list stops)."""
    mask = 0
    for name in names:
        mask |= 1 << KERNEL_NAMES.index(name)
    rc = lib().bi_hip_profile_kinds(mask)
    if rc:
        raise EngineError(rc, "bi_hip_profile_kinds")


def profile_read() -> dict:
    """Per kernel class: {'ms': total event ms, 'bytes': algorithmic bytes, 'launches': count}."""
    arr = (KStat * len(KERNEL_NAMES))()
    rc = lib().bi_hip_profile_read(arr, len(KERNEL_NAMES))
    if rc:
        raise EngineError(rc, "bi_hip_profile_read")
    return {name: {"ms": arr[i].ms, "bytes": arr[i].bytes, "launches": int(arr[i].launches)}
            for i, name in enumerate(KERNEL_NAMES)}
