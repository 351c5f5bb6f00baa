"""bi_hip_mul_device_noalloc: the allocation-free variant of bi_hip_mul_device.

Called through the C ABI and checked limb for limb against bi_hip_mul_device (and the oracle) on
two calls in a row, the second with a smaller input, then a square; sizes beyond its fixed
scratch are rejected with BI_HIP_ESCRATCH, a process that never called bi_hip_init gets
BI_HIP_ENOINIT, and device free memory does not move across its calls.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

from bigint_amd import engine, hip
from bigint_amd import limbs as L
from oracle import oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _free_bytes() -> int:
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.rt().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


class _Case:
    """Device copies of one product's operands and an output buffer, allocated up front."""

    def __init__(self, a, b, square):
        self.a, self.b, self.square = a, (a if square else b), square
        self.da = hip.Buffer.from_array(a)
        self.db = self.da if square else hip.Buffer.from_array(self.b)
        self.cap = a.size + self.b.size
        self.do = hip.Buffer(4 * self.cap)

    def run(self, fn):
        """The product through fn (a C-ABI entry point); returns its la + lb limbs."""
        self.do.zero()
        rc = fn(self.da.ptr, self.a.size, None if self.square else self.db.ptr,
                0 if self.square else self.b.size, int(self.square), self.do.ptr, self.cap, None)
        assert rc == 0, engine.strerror(rc)
        assert engine.lib().bi_hip_synchronize() == 0
        return self.do.download(np.uint32, self.cap)

    def free(self):
        for x in {id(self.da): self.da, id(self.db): self.db, id(self.do): self.do}.values():
            x.free()


def test_noalloc_matches_mul_device():
    engine.init(0)
    lib = engine.lib()
    big_a = L.random_operand(1 << 24, 301)        # 2^19 words each: a 2^20-point transform, the maximum
    big_b = L.random_operand((1 << 24) - 5, 302)
    assert engine.transform_log2(big_a.size, big_b.size) == 20
    small_a = L.random_operand(100_003, 303)
    small_b = L.random_operand(70_001, 304)
    sq = L.random_operand(3_000_000, 305)
    cases = [_Case(big_a, big_b, False), _Case(small_a, small_b, False), _Case(sq, None, True)]
    ref = [c.run(lib.bi_hip_mul_device) for c in cases]
    cases[1].run(lib.bi_hip_mul_device_noalloc)  # first launch of its kernels
    free0 = _free_bytes()
    # two calls in a row, the second smaller, then a square
    got = [c.run(lib.bi_hip_mul_device_noalloc) for c in cases]
    # the large product once more right after the small ones: stale scratch from the last call
    # must not leak into it
    again = cases[0].run(lib.bi_hip_mul_device_noalloc)
    assert _free_bytes() == free0  # nothing allocated by the calls
    for c, r, g in zip(cases, ref, got):
        assert np.array_equal(g, r), (c.a.size, c.square)
        if c.a.size < 200_000:
            assert np.array_equal(L.strip(g), oracle.mul_limbs(c.a, None if c.square else c.b))
    assert np.array_equal(again, ref[0])
    for c in cases:
        c.free()
    # beyond the scratch: a 2^21-point transform
    a = L.random_operand(1 << 25, 306)
    da, do = hip.Buffer.from_array(a), hip.Buffer(8 * a.size)
    rc = lib.bi_hip_mul_device_noalloc(da.ptr, a.size, None, 0, 1, do.ptr, 2 * a.size, None)
    assert rc == engine.ESCRATCH
    assert b"scratch" in lib.bi_hip_strerror(rc)
    da.free()
    do.free()


def test_noalloc_before_init_is_rejected():
    """In a fresh process without bi_hip_init the call returns BI_HIP_ENOINIT (and touches no
    device memory: the pointers below are never dereferenced)."""
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from bigint_amd import engine\n"
            "rc = engine.lib().bi_hip_mul_device_noalloc(8, 1, 8, 1, 0, 8, 2, None)\n"
            "print(rc)\n") % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert int(out.stdout.strip().splitlines()[-1]) == engine.ENOINIT
