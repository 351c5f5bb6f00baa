"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py import this module,
and only as the checker.  The product (bigint_amd/) never calls it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_MT = None
_CHK = None

u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
szp = ctypes.POINTER(ctypes.c_size_t)


def build(force: bool = False) -> None:
    """Compile liboracle.so and libcpu_mt.so with the recipe in oracle/Makefile."""
    targets = [os.path.join(HERE, n) for n in ("liboracle.so", "libcpu_mt.so", "libbatch_check.so")]
    if force or not all(os.path.exists(t) for t in targets):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"] + ([] if not force else ["all"]),
                       check=True)


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.ora_mont_mul.restype = ctypes.c_uint32
        L.ora_mont_mul.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32]
        L.ora_garner.argtypes = [ctypes.c_uint32] * 3 + [u64p, u64p]
        L.ora_imm_ntt.argtypes = [u32p, ctypes.c_size_t, ctypes.c_int]
        L.ora_get_constants.argtypes = [u32p] * 7
        L.ora_ntt_mul_u64.argtypes = [u64p, ctypes.c_size_t, u64p, ctypes.c_size_t, ctypes.c_int, u64p, szp]
        L.ora_mul_u64.argtypes = L.ora_ntt_mul_u64.argtypes
        L.ora_brute_mul_u64.argtypes = [u64p, ctypes.c_size_t, u64p, ctypes.c_size_t, u64p, szp]
        L.ora_mul_digits.argtypes = [u32p, ctypes.c_size_t, u32p, ctypes.c_size_t, ctypes.c_int,
                                     ctypes.c_size_t, u32p, szp, ctypes.POINTER(ctypes.c_int64)]
        L.ora_mul_limbs.argtypes = [u32p, ctypes.c_size_t, u32p, ctypes.c_size_t, ctypes.c_int, u32p, szp]
        L.ora_init()
        _LIB = L
    return _LIB


def _p32(a: np.ndarray):
    return a.ctypes.data_as(u32p)


def _p64(a: np.ndarray):
    return a.ctypes.data_as(u64p)


def constants() -> dict:
    L = lib()
    bufs = [np.zeros(3, np.uint32) for _ in range(5)]
    steps = np.zeros((2, 26, 3), np.uint32)
    inv2 = np.zeros((26, 3), np.uint32)
    L.ora_get_constants(*[_p32(b) for b in bufs], _p32(steps), _p32(inv2))
    names = ["P", "G", "neg_inv_P", "R_sq", "vec_one"]
    d = {k: [int(x) for x in b] for k, b in zip(names, bufs)}
    d["steps"] = steps
    d["inv_pow_of_two"] = inv2
    return d


def mont_mul(I: int, a: int, b: int) -> int:
    return int(lib().ora_mont_mul(I, a, b))


def garner(c0: int, c1: int, c2: int) -> int:
    lo, hi = ctypes.c_uint64(), ctypes.c_uint64()
    lib().ora_garner(c0, c1, c2, ctypes.byref(lo), ctypes.byref(hi))
    return (hi.value << 64) | lo.value


def imm_ntt(v: np.ndarray, rev: bool) -> np.ndarray:
    """In-place interleaved 3-prime transform of a copy of ``v`` (length 3n)."""
    v = np.ascontiguousarray(v, dtype=np.uint32).copy()
    lib().ora_imm_ntt(_p32(v), v.size // 3, int(rev))
    return v


def ntt_mul_u64(a: np.ndarray, b: np.ndarray | None) -> np.ndarray:
    """ntt::mul on u64 words; b=None selects the squaring branch."""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    same = b is None
    bb = a if same else np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros(a.size + bb.size + 1, np.uint64)
    n = ctypes.c_size_t()
    rc = lib().ora_ntt_mul_u64(_p64(a), a.size, _p64(bb), bb.size, int(same), _p64(out), ctypes.byref(n))
    if rc:
        raise MemoryError("oracle ntt_mul failed")
    return out[: n.value]


def mul_digits(a: np.ndarray, b: np.ndarray | None, precision: int = 0) -> tuple[np.ndarray, int]:
    """mul_digits on 28-bit limbs; b=None means the same object (squaring).  Returns the limbs
    and the amount the caller's point position is reduced by."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    same = b is None
    bb = a if same else np.ascontiguousarray(b, dtype=np.uint32)
    out = np.zeros(a.size + bb.size + 4, np.uint32)
    n = ctypes.c_size_t()
    pd = ctypes.c_int64()
    rc = lib().ora_mul_digits(_p32(a), a.size, _p32(bb), bb.size, int(same), precision, _p32(out),
                              ctypes.byref(n), ctypes.byref(pd))
    if rc:
        raise RuntimeError(f"oracle mul_digits failed rc={rc}")
    return out[: n.value].copy(), pd.value


def mul_limbs(a: np.ndarray, b: np.ndarray | None) -> np.ndarray:
    return mul_digits(a, b, 0)[0]


# ---- threaded AVX2 restatement (CPU baseline) -------------------------------------------------

def mt_lib() -> ctypes.CDLL:
    global _MT
    if _MT is None:
        path = os.path.join(HERE, "libcpu_mt.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.cpu_init_thread_pool.argtypes = [ctypes.c_uint32]
        L.cpu_mul_digits.restype = ctypes.c_int
        L.cpu_mul_digits.argtypes = [u32p, ctypes.c_size_t, u32p, ctypes.c_size_t, ctypes.c_int,
                                     u32p, szp]
        L.cpu_pool_threads.restype = ctypes.c_uint32
        L.cpu_route.restype = ctypes.c_int
        L.cpu_route.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
        _MT = L
    return _MT


# ---- large-batch checker (batch_check.c) ------------------------------------------------------

def chk_lib() -> ctypes.CDLL:
    global _CHK
    if _CHK is None:
        path = os.path.join(HERE, "libbatch_check.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.chk_random_operand.restype = ctypes.c_size_t
        L.chk_random_operand.argtypes = [ctypes.c_uint64, ctypes.c_uint64, u32p]
        L.chk_residues.restype = ctypes.c_int
        L.chk_residues.argtypes = [u32p, u64p, u64p, ctypes.c_size_t, ctypes.c_uint32, u32p, ctypes.c_int]
        _CHK = L
    return _CHK


def random_operands(bits: list[int], seeds: list[int]) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """limbs.random_operand(bits[i], seeds[i]) for every i, concatenated into one pool; returns
    (pool, offsets, lengths) in limbs."""
    lens = np.array([-(-b // 28) for b in bits], np.uint64)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint64)
    pool = np.zeros(int(lens.sum()), np.uint32)
    L = chk_lib()
    base = pool.ctypes.data
    for b, s, o in zip(bits, seeds, offs):
        L.chk_random_operand(int(b), int(s) & ((1 << 64) - 1),
                             ctypes.cast(base + 4 * int(o), u32p))
    return pool, offs, lens


def residues(pool: np.ndarray, offs: np.ndarray, lens: np.ndarray, q: int, threads: int = 16) -> np.ndarray:
    """Value of each segment pool[offs[i] : offs[i] + lens[i]] (28-bit limbs) mod q."""
    pool = np.ascontiguousarray(pool, dtype=np.uint32)
    offs = np.ascontiguousarray(offs, dtype=np.uint64)
    lens = np.ascontiguousarray(lens, dtype=np.uint64)
    assert offs.size == lens.size and (lens.size == 0 or int((offs + lens).max()) <= pool.size)
    out = np.zeros(offs.size, np.uint32)
    if chk_lib().chk_residues(_p32(pool), _p64(offs), _p64(lens), offs.size, q, _p32(out), threads):
        raise ValueError("modulus out of range")
    return out.astype(np.int64)
