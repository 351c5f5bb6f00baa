"""Pins the CPU oracle (oracle/ntt_oracle.c) before anything is checked against it.

Each check is chosen so that a plausible mistake in the restatement (a wrong root, a dropped
Montgomery factor, a transposed Garner operand, a wrong trim weight) fails:
  - constants against their defining properties (order of roots, inverses mod 2^32 and mod P);
  - Montgomery REDC against Python pow() at edge and random values;
  - the transform against a direct O(n^2) DFT mod P, and inverse(forward(x)) == x;
  - Garner against CRT computed with Python ints;
  - products against CPython's int multiplication (golden fixtures + on-the-fly) and closed
    forms such as (2^a - 1)(2^b - 1), the reference's own worst case
    (tests/measure_digit_bits.cpp:65-85);
  - mul_digits precision/zero-stripping rules against an independent Python model of
    include/bigint/mul.h:223-318.
"""
import hashlib
import json
import os
import random

import numpy as np
import pytest

from bigint_amd import limbs as L
from oracle import oracle

P = [469762049, 2013265921, 2281701377]
G = [3, 31, 3]
R = 1 << 32
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "products.json")


def is_prime(n: int) -> bool:
    if n < 2:
        return False
    for a in [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]:
        if n % a == 0:
            return n == a
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def test_constants():
    c = oracle.constants()
    assert c["P"] == P and c["G"] == G
    for I, p in enumerate(P):
        assert is_prime(p)
        # G generates the full 2-power part: G^((p-1)/2) = -1
        assert pow(G[I], (p - 1) // 2, p) == p - 1
        assert (c["neg_inv_P"][I] * p) % R == R - 1
        assert c["R_sq"][I] == (R * R) % p
        assert c["vec_one"][I] == R % p
        rinv = pow(R, -1, p)
        for level in range(1, 27):
            w = int(c["steps"][0][level - 1][I]) * rinv % p  # leave Montgomery form
            wi = int(c["steps"][1][level - 1][I]) * rinv % p
            assert pow(w, 1 << level, p) == 1
            assert pow(w, 1 << (level - 1), p) == p - 1     # primitive 2^level-th root
            assert w * wi % p == 1
            assert int(c["inv_pow_of_two"][level - 1][I]) * rinv % p == pow(2, -level, p)
    # max transform length is 2^26: the smallest 2-adic valuation of p-1
    assert min(((p - 1) & -(p - 1)).bit_length() - 1 for p in P) == 26


def test_mont_mul():
    rng = random.Random(5)
    for I, p in enumerate(P):
        rinv = pow(R, -1, p)
        vals = [0, 1, 2, p - 1, p - 2, p // 2] + [rng.randrange(p) for _ in range(200)]
        for a in vals:
            for b in vals[:12]:
                assert oracle.mont_mul(I, a, b) == a * b * rinv % p


def test_garner():
    rng = random.Random(6)
    M = P[0] * P[1] * P[2]
    for _ in range(2000):
        x = rng.randrange(M)
        assert oracle.garner(x % P[0], x % P[1], x % P[2]) == x
    for x in [0, 1, M - 1, P[0], P[0] * P[1], P[0] * P[1] - 1]:
        assert oracle.garner(x % P[0], x % P[1], x % P[2]) == x


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64])
def test_transform_is_dft(n):
    rng = np.random.default_rng(n)
    x = np.stack([rng.integers(0, p, n, dtype=np.uint64) for p in P], axis=1).astype(np.uint32)
    y = oracle.imm_ntt(x.reshape(-1), False).reshape(n, 3)
    for I, p in enumerate(P):
        w = pow(G[I], (p - 1) // n, p)
        xs = [int(v) for v in x[:, I]]
        dft = [sum(xs[j] * pow(w, j * k, p) for j in range(n)) % p for k in range(n)]
        assert [int(v) for v in y[:, I]] == dft
    back = oracle.imm_ntt(y.reshape(-1), True).reshape(n, 3)
    assert np.array_equal(back, x)


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def _operands(case):
    a = L.random_operand(case["bits_a"], case["seed_a"])
    b = None if case["bits_b"] is None else L.random_operand(case["bits_b"], case["seed_b"])
    return a, b


def _sha(limbs):
    return hashlib.sha256(np.ascontiguousarray(limbs, dtype="<u4").tobytes()).hexdigest()


@pytest.mark.parametrize("case", _golden(), ids=lambda c: c["name"])
def test_golden_products(case):
    a, b = _operands(case)
    p = oracle.mul_limbs(a, b)
    assert p.size == case["n_limbs"]
    assert _sha(p) == case["sha256"]
    if "limbs" in case:
        assert [int(x) for x in p] == case["limbs"]


def test_ntt_path_directly():
    """ntt::mul on u64 words, bypassing the dispatcher (which sends tiny inputs to brute)."""
    rng = random.Random(9)
    for na, nb in [(1, 1), (1, 3), (2, 2), (5, 300), (64, 64), (129, 1000)]:
        a = [rng.getrandbits(64) | 1 for _ in range(na)]
        b = [rng.getrandbits(64) | 1 for _ in range(nb)]
        A, B = np.array(a, np.uint64), np.array(b, np.uint64)
        ia = sum(x << (64 * i) for i, x in enumerate(a))
        ib = sum(x << (64 * i) for i, x in enumerate(b))
        got = oracle.ntt_mul_u64(A, B)
        assert sum(int(x) << (64 * i) for i, x in enumerate(got)) == ia * ib
        got = oracle.ntt_mul_u64(A, None)
        assert sum(int(x) << (64 * i) for i, x in enumerate(got)) == ia * ia


@pytest.mark.parametrize("a,b", [(28, 28), (56, 28), (100, 3), (1000, 1000), (28 * 4096, 28 * 4096),
                                 (30000, 17), (12345, 54321), (28 * 70000, 28 * 70000)])
def test_closed_form_all_ones(a, b):
    x, y = (1 << a) - 1, (1 << b) - 1
    p = oracle.mul_limbs(L.int_to_limbs(x), L.int_to_limbs(y))
    assert L.limbs_to_int(p) == (1 << (a + b)) - (1 << a) - (1 << b) + 1
    sq = oracle.mul_limbs(L.int_to_limbs(x), None)
    assert L.limbs_to_int(sq) == x * x


def test_shifts_and_identity():
    x = L.random_operand(20000, 77)
    ix = L.limbs_to_int(x)
    for k in [0, 1, 27, 28, 29, 1000, 28 * 500]:
        p = oracle.mul_limbs(x, L.int_to_limbs(1 << k))
        assert L.limbs_to_int(p) == ix << k
    assert np.array_equal(oracle.mul_limbs(x, L.int_to_limbs(1)), x)


def test_commutative_and_distributive():
    a, b, c = (L.random_operand(b, s) for b, s in [(9000, 1), (15000, 2), (7000, 3)])
    ab, ba = oracle.mul_limbs(a, b), oracle.mul_limbs(b, a)
    assert np.array_equal(ab, ba)
    ia, ib, ic = (L.limbs_to_int(v) for v in (a, b, c))
    lhs = L.limbs_to_int(oracle.mul_limbs(L.int_to_limbs(ia + ib), c))
    assert lhs == L.limbs_to_int(oracle.mul_limbs(a, c)) + L.limbs_to_int(oracle.mul_limbs(b, c))


def model_mul_digits(a, b, precision):
    """Independent model of mul.h:223-318 on Python ints and lists."""
    a, b = list(a), list(b)
    oa = 0 if precision == 0 or len(a) <= precision else len(a) - precision
    ob = 0 if precision == 0 or len(b) <= precision else len(b) - precision
    za = next(i for i in range(oa, len(a)) if a[i]) - oa
    zb = next(i for i in range(ob, len(b)) if b[i]) - ob
    oa, ob = oa + za, ob + zb
    va = L.limbs_to_int(np.array(a[oa:], np.uint32))
    vb = L.limbs_to_int(np.array(b[ob:], np.uint32))
    res = [int(v) for v in L.int_to_limbs(va * vb)]
    if precision:
        ores = 0 if len(res) <= precision else len(res) - precision
        while res[ores] == 0:
            ores += 1
        return res[ores:], oa + ob + ores
    return [0] * (za + zb) + res, 0


def test_mul_digits_precision_rules():
    rng = random.Random(11)
    for trial in range(60):
        na, nb = rng.randrange(1, 400), rng.randrange(1, 400)
        a = [rng.randrange(1 << 28) for _ in range(na)]
        b = [rng.randrange(1 << 28) for _ in range(nb)]
        a[-1] |= 1
        b[-1] |= 1
        # sprinkle zero limbs, including trailing ones, to exercise the stripping rules
        for v in (a, b):
            for _ in range(rng.randrange(0, 5)):
                v[rng.randrange(len(v))] = 0
            v[-1] |= 1
            if rng.random() < 0.5:
                z = rng.randrange(1, 6)
                v[:z] = [0] * min(z, len(v) - 1)
        prec = rng.choice([0, 1, 2, 7, 50, 300, 1000])
        got, pd = oracle.mul_digits(np.array(a, np.uint32), np.array(b, np.uint32), prec)
        exp, epd = model_mul_digits(a, b, prec)
        assert [int(x) for x in got] == exp, (na, nb, prec)
        assert pd == epd
    # squaring through the same object matches the general path
    a = np.array([0, 0, 5, 7, 1], np.uint32)
    sq, _ = oracle.mul_digits(a, None, 0)
    gen, _ = oracle.mul_digits(a, a.copy(), 0)
    assert np.array_equal(sq, gen)
