"""The threaded AVX2 restatement used as the CPU baseline (oracle/libcpu_mt.so) computes the
same products as the pinned scalar oracle and CPython ints."""
import ctypes

import numpy as np
import pytest

from bigint_amd import limbs as L
from oracle import oracle


def mt_mul(a, b):
    M = oracle.mt_lib()
    M.cpu_init_thread_pool(4)
    same = b is None
    bb = a if same else b
    out = np.zeros(a.size + bb.size + 4, np.uint32)
    n = ctypes.c_size_t()
    p = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))  # noqa: E731
    M.cpu_mul_digits(p(a), a.size, p(bb), bb.size, int(same), p(out), ctypes.byref(n))
    return out[: n.value]


@pytest.mark.parametrize("ba,bb", [(2000, 3000), (1 << 16, 1 << 16), (1 << 20, 1 << 12),
                                   (1 << 21, 1 << 21), (28 * 3000, 28 * 70000)])
def test_matches_oracle(ba, bb):
    a = L.random_operand(ba, 1)
    b = L.random_operand(bb, 2)
    assert np.array_equal(mt_mul(a, b), oracle.mul_limbs(a, b))
    assert np.array_equal(mt_mul(a, None), oracle.mul_limbs(a, None))


def test_trailing_zero_limbs_and_all_ones():
    x = (1 << (28 * 5000)) - 1
    a = L.int_to_limbs(x)
    b = np.concatenate([np.zeros(7, np.uint32), L.random_operand(40000, 3)])
    assert L.limbs_to_int(mt_mul(a, b)) == x * L.limbs_to_int(b)
