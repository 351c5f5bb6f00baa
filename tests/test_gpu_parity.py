"""Parity of the HIP engine (through the C ABI) with the oracle and the golden products.

Bar: bit-exact limbs.  Sizes the oracle finishes in seconds are compared limb for limb with the
oracle or with the committed CPython-int fixtures; the full 2^26 and 2^28 workloads are checked
through size-independent properties: residues modulo six primes below 2^31
((a mod q)(b mod q) == (a*b) mod q), commutativity and square == general product.
"""
import hashlib
import json
import os
import random

import numpy as np
import pytest

from bigint_amd import engine, hip
from bigint_amd import limbs as L
from oracle import oracle

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "products.json")


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init(0)


def sha(limbs):
    return hashlib.sha256(np.ascontiguousarray(limbs, dtype="<u4").tobytes()).hexdigest()


def check_residues(a, b, p):
    for q in L.RESIDUE_PRIMES:
        ra = L.residue_mod(a, q)
        rb = ra if b is None else L.residue_mod(b, q)
        assert L.residue_mod(p, q) == ra * rb % q, q


with open(GOLDEN) as _f:
    _CASES = json.load(_f)["cases"]


@pytest.mark.parametrize("case", _CASES, ids=lambda c: c["name"])
def test_golden(case):
    a = L.random_operand(case["bits_a"], case["seed_a"])
    b = None if case["bits_b"] is None else L.random_operand(case["bits_b"], case["seed_b"])
    p = engine.mul(a, b)
    assert p.size == case["n_limbs"]
    assert sha(p) == case["sha256"]


def test_every_small_length_against_oracle():
    """All la, lb in 1..40 plus lengths around every transform-length boundary up to 2^13."""
    rng = np.random.default_rng(3)
    sizes = [(la, lb) for la in range(1, 41) for lb in range(1, 41, 3)]
    for k in range(4, 14):
        for d in (-1, 0, 1):
            tot = (1 << k) + 1 + d
            sizes += [(tot // 2, tot - tot // 2), (1, tot - 1), (tot - 3, 3)]
    for la, lb in sizes:
        if la < 1 or lb < 1:
            continue
        a = rng.integers(0, 1 << 28, la, dtype=np.uint32)
        b = rng.integers(0, 1 << 28, lb, dtype=np.uint32)
        a[-1] |= 1
        b[-1] |= 1
        got = engine.mul(a, b)
        exp = oracle.mul_limbs(a, b)
        assert np.array_equal(got, exp), (la, lb)


@pytest.mark.parametrize("bits", [28, 29, 1000, 28 * 2048, 28 * 2049, 28 * 4096, 123457, 1 << 20])
def test_squares_against_oracle(bits):
    a = L.random_operand(bits, 100 + bits)
    assert np.array_equal(engine.mul(a, None), oracle.mul_limbs(a, None))
    assert np.array_equal(engine.mul(a, None), engine.mul(a, a.copy()))


@pytest.mark.parametrize("a_bits,b_bits", [(28, 28), (28 * 8, 28 * 8), (28 * 2048, 28 * 2048),
                                           (28 * 4096, 28 * 4096), (28 * 70000, 28 * 70000),
                                           (28 * 300000, 28 * 5000), (1 << 22, 1 << 22)])
def test_all_ones_worst_case(a_bits, b_bits):
    """(2^a - 1)(2^b - 1): every limb 2^28-1 maximises every convolution coefficient."""
    x, y = (1 << a_bits) - 1, (1 << b_bits) - 1
    p = engine.mul(L.int_to_limbs(x), L.int_to_limbs(y))
    assert L.limbs_to_int(p) == (1 << (a_bits + b_bits)) - (1 << a_bits) - (1 << b_bits) + 1


def test_random_mid_sizes_against_oracle():
    rng = random.Random(17)
    for _ in range(25):
        la = int(2 ** rng.uniform(4, 17))
        lb = int(2 ** rng.uniform(0, 17))
        a = L.random_operand(28 * la - rng.randrange(28), rng.getrandbits(64))
        b = L.random_operand(28 * lb - rng.randrange(28), rng.getrandbits(64))
        assert np.array_equal(engine.mul(a, b), oracle.mul_limbs(a, b)), (la, lb)


@pytest.mark.parametrize("a_bits,b_bits", [(1 << 26, 1 << 26), (1 << 28, 1 << 20), (1 << 28, 1 << 28)])
def test_large_workloads_residues(a_bits, b_bits):
    a = L.random_operand(a_bits, L.SEED)
    b = L.random_operand(b_bits, L.SEED + 1)
    p = engine.mul(a, b)
    assert p.size in (a.size + b.size, a.size + b.size - 1)
    assert p[-1] != 0
    check_residues(a, b, p)
    if a_bits == b_bits == 1 << 26:
        assert np.array_equal(p, engine.mul(b, a))
        s = engine.mul(a, None)
        check_residues(a, None, s)


def test_max_transform_length():
    """2^25 + 1 words per operand on a 2^26-point transform (the top coefficient wraps onto
    coefficient 0 and is separated again): the largest product the three primes allow."""
    la = lb = 38347923
    a = L.random_operand(28 * la, 5)
    b = L.random_operand(28 * lb - 3, 6)
    assert engine.transform_log2(la, lb) == 26
    p = engine.mul(a, b)
    check_residues(a, b, p)


def test_error_paths():
    a = np.array([1, 2, 3], np.uint32)
    with pytest.raises(engine.EngineError) as e:
        engine.mul(np.array([1, 0], np.uint32), a)  # top limb zero
    assert e.value.code == engine.EINVAL
    # a product of 2^32 - 8192 limbs or more is rejected on the host before any device work (the
    # lengths are checked before the pointers are used)
    buf = hip.Buffer(64)
    for la, lb in [(1 << 31, 1 << 31), ((1 << 32) - 8192, 1), (1, (1 << 32) + 5)]:
        with pytest.raises(engine.EngineError) as e:
            engine.mul_device(buf.ptr, la, buf.ptr, lb, False, buf.ptr, la + lb, None)
        assert e.value.code == engine.ETOOBIG, (la, lb)


def test_mul_top_matches_full_product():
    """bi_hip_mul_top (the precision-truncated route of mul_digits) returns exactly the top limbs
    of the oracle's product and the whole product's length, including products whose top limb
    is zero and squares."""
    rng = random.Random(21)
    cases = [(L.random_operand(56000, 1), L.random_operand(84000, 2)),
             (L.random_operand(200000, 3), None),
             (np.array([1] + [0] * 2999 + [1], np.uint32), np.array([5] * 1999 + [1], np.uint32)),
             (np.full(3000, (1 << 28) - 1, np.uint32), np.full(2500, (1 << 28) - 1, np.uint32))]
    for a, b in cases:
        full = oracle.mul_limbs(a, b)
        m = a.size + (a.size if b is None else b.size)
        padded = np.zeros(m, np.uint32)
        padded[: full.size] = full
        for keep in [1, 2, 7, rng.randrange(1, m), m - 1, m, m + 5]:
            top, n = engine.mul_top(a, b, keep)
            k = min(keep, m)
            assert n == full.size, keep
            assert np.array_equal(top[:k], padded[m - k:]), keep
    with pytest.raises(engine.EngineError) as e:
        engine.mul_top(cases[0][0], cases[0][1], 0)
    assert e.value.code == engine.EINVAL


def test_device_add_batch_against_python_ints():
    """bi_hip_add_batch_device: sums and differences of device limb arrays, bit-exact against
    Python ints, with carries and borrows running through whole 4096-limb tiles, equal
    magnitudes (zero), both subtraction directions, unequal lengths and zero top limbs."""
    from bigint_amd import hip
    rng = random.Random(33)
    M = (1 << 28) - 1

    def limbs_of(v, n):
        return np.array([(v >> (28 * i)) & M for i in range(n)], np.uint32)

    def val(a):
        return sum(int(x) << (28 * i) for i, x in enumerate(a))

    cases = []
    for lx, ly in [(1, 1), (5, 3), (4096, 4096), (4097, 1), (1, 9000), (12289, 12288), (30000, 29999)]:
        cases.append((L.random_operand(28 * lx, rng.getrandbits(63))[:lx],
                      L.random_operand(28 * ly, rng.getrandbits(63))[:ly]))
    ones = np.full(9000, M, np.uint32)
    cases.append((ones, np.array([1], np.uint32)))            # carry through three tiles
    big = np.zeros(9000, np.uint32)
    big[-1] = 1
    cases.append((big, np.array([1], np.uint32)))             # borrow through three tiles
    same = L.random_operand(28 * 5000, 7)
    cases.append((same, same.copy()))                         # equal magnitudes
    cases.append((np.append(same, np.uint32(0)), same.copy()))  # zero top limb
    bufs, jobs, expect = [], [], []
    for x, y in cases:
        for sub in (False, True):
            dx, dy = hip.Buffer.from_array(x), hip.Buffer.from_array(y)
            cap = max(x.size, y.size) + 1
            do = hip.Buffer(4 * cap)
            bufs += [dx, dy, do]
            jobs.append((dx.ptr, x.size, dy.ptr, y.size, sub, do.ptr, cap))
            vx, vy = val(x), val(y)
            expect.append((do, vx + vy if not sub else abs(vx - vy), sub and vy > vx))
    res = engine.add_batch_device(jobs)
    for (do, want, swapped), (sw, n) in zip(expect, res):
        got = do.download(np.uint32, max(n, 1))
        assert val(got[:n]) == want
        assert n == (want.bit_length() + 27) // 28  # 0 for zero
        assert sw == swapped


def test_batch_beyond_one_launch_row_limit():
    """A grouped batch of 70,000 products (one transform length, so one group): its launches put
    one job per grid row and split past 65,535 rows.  Every product is checked modulo two primes,
    a sample limb for limb against the oracle."""
    rng = np.random.default_rng(41)
    count = 70000
    a = rng.integers(1, 1 << 28, size=(count, 60), dtype=np.uint32)
    b = rng.integers(1, 1 << 28, size=(count, 50), dtype=np.uint32)
    pairs = [(a[i], b[i]) for i in range(count)]
    outs = engine.mul_batch(pairs, engine.PLAN_TABLE)
    for q in L.RESIDUE_PRIMES[:2]:
        for i in range(count):
            assert L.residue_mod(outs[i], q) == L.residue_mod(a[i], q) * L.residue_mod(b[i], q) % q, i
    for i in [0, 1, 65534, 65535, 65536, count - 1]:
        assert np.array_equal(outs[i], oracle.mul_limbs(a[i], b[i])), i


def test_batch_matches_single():
    rng = random.Random(4)
    pairs = []
    for _ in range(12):
        a = L.random_operand(rng.randrange(28, 200000), rng.getrandbits(63))
        b = None if rng.random() < 0.3 else L.random_operand(rng.randrange(28, 200000), rng.getrandbits(63))
        pairs.append((a, b))
    outs = engine.mul_batch(pairs)
    for (a, b), p in zip(pairs, outs):
        assert np.array_equal(p, oracle.mul_limbs(a, b))


@pytest.mark.parametrize("mode", [engine.PLAN_FIXED, engine.PLAN_TABLE])
def test_batch_grouped_matches_oracle(mode):
    """Many products of mixed transform lengths in one grouped launch sequence."""
    rng = random.Random(40 + mode)
    pairs = []
    for i in range(60):
        bits_a = int(2 ** rng.uniform(5, 18))
        a = L.random_operand(bits_a, rng.getrandbits(63))
        if i % 7 == 0:
            b = None
        else:
            b = L.random_operand(int(2 ** rng.uniform(5, 18)), rng.getrandbits(63))
        pairs.append((a, b))
    # equal lengths too, so some groups have many jobs
    for i in range(20):
        pairs.append((L.random_operand(28 * 3000, 1000 + i), L.random_operand(28 * 2000, 2000 + i)))
    outs = engine.mul_batch(pairs, mode)
    for (a, b), p in zip(pairs, outs):
        assert np.array_equal(p, oracle.mul_limbs(a, b))


@pytest.mark.parametrize("a_bits,b_bits", [(28 * 20000, 28 * 300), (1 << 22, 1 << 14), (1 << 16, 1 << 21),
                                           (28 * 4096 * 9 + 5, 28 * 2049)])
def test_chunked_unbalanced_against_oracle(a_bits, b_bits):
    """The grouped path cuts the longer operand into chunks; partial products overlap-add
    before the carry scan."""
    a = L.random_operand(a_bits, 71)
    b = L.random_operand(b_bits, 72)
    (p,) = engine.mul_batch([(a, b)], engine.PLAN_TABLE)
    assert np.array_equal(p, oracle.mul_limbs(a, b))


def test_chunked_all_ones_worst_case():
    x, y = (1 << (28 * 50000)) - 1, (1 << (28 * 3000)) - 1
    (p,) = engine.mul_batch([(L.int_to_limbs(x), L.int_to_limbs(y))], engine.PLAN_TABLE)
    assert L.limbs_to_int(p) == x * y


def test_chunked_2e28_by_2e20_residues():
    a = L.random_operand(1 << 28, L.SEED)
    b = L.random_operand(1 << 20, L.SEED + 1)
    (p,) = engine.mul_batch([(a, b)], engine.PLAN_TABLE)
    assert np.array_equal(p, engine.mul(a, b))
    check_residues(a, b, p)


def _limbs_for_words(w: int) -> int:
    """Largest limb count whose 32-bit word count is w."""
    la = (8 * w) // 7
    while (la * 7 + 7) // 8 > w:
        la -= 1
    while ((la + 1) * 7 + 7) // 8 == w:
        la += 1
    return la


@pytest.mark.parametrize("wa,wb", [(2049, 2049), (4097, 4097), (3000, 1098), (65537, 65537),
                                   (1 << 19, (1 << 19) + 2), ((1 << 21) + 1, (1 << 21) + 1)])
def test_wrapped_top_coefficient(wa, wb):
    """wa + wb - 1 = n + 1: the top coefficient wraps onto coefficient 0 of the cyclic
    convolution and Garner separates them.  All-ones operands make both top words (and every
    coefficient) maximal; random operands and the square cover the rest."""
    la, lb = _limbs_for_words(wa), _limbs_for_words(wb)
    n = 1 << engine.transform_log2(la, lb)
    assert wa + wb - 1 == n + 1
    x, y = (1 << (28 * la)) - 1, (1 << (28 * lb)) - 1
    p = engine.mul(L.int_to_limbs(x), L.int_to_limbs(y))
    assert L.limbs_to_int(p) == x * y
    a = L.random_operand(28 * la, wa)
    b = L.random_operand(28 * lb, wb)
    p = engine.mul(a, b)
    if la <= 70000:
        assert np.array_equal(p, oracle.mul_limbs(a, b))
    else:
        check_residues(a, b, p)
    if wa == wb:
        check_residues(a, None, engine.mul(a, None))


def test_carry_runs_across_tiles():
    """a * (2^32 + 1) is a + (a << 32): with a's words alternating so that a_j + a_(j-1) =
    2^32 - 1, one carry from word 2 runs through three whole carry tiles (3584 words each).
    Those tiles cannot resolve their carry-in locally: they are written with carry-in 0 and the
    carry fix-up composes the tile maps back to the tile where the carry was born."""
    W = 3 * 3584 + 100
    words = [(1 << 32) - 1, 5] + [((1 << 32) - 6) if j % 2 == 0 else 5 for j in range(2, W)]
    a = sum(w << (32 * j) for j, w in enumerate(words))
    b = (1 << 32) + 1
    p = engine.mul(L.int_to_limbs(a), L.int_to_limbs(b))
    assert L.limbs_to_int(p) == a * b
    # and the square of such a run, for carries born inside the product
    p2 = engine.mul(L.int_to_limbs(a), None)
    assert L.limbs_to_int(p2) == a * a


def test_carry_runs_across_a_thousand_tiles():
    """The same construction over 1,100 carry tiles (3.94 M words, 126 Mbit): every tile after
    the first is written with an assumed carry-in and fixed, so the fix-up scans its tile maps in
    two chunks of 1024 and ripples a carry through every tile."""
    W = 1100 * 3584 + 100
    words = np.full(W, 5, dtype=np.uint64)
    words[2::2] = (1 << 32) - 6
    words[0] = (1 << 32) - 1
    a = int.from_bytes(words.astype("<u4").tobytes(), "little")
    b = (1 << 32) + 1
    p = engine.mul(L.int_to_limbs(a), L.int_to_limbs(b))
    assert L.limbs_to_int(p) == a * b


def check_residues3(a, b, p):
    for q in L.RESIDUE_PRIMES[:3]:
        ra = L.residue_mod(a, q)
        rb = ra if b is None else L.residue_mod(b, q)
        assert L.residue_mod(p, q) == ra * rb % q, q


def test_beyond_one_transform_both_operands_chunked():
    """50.75 M words (1.62 Gbit) per operand: beyond one 2^26-point transform, and too long for
    any chunk plan of one operand, so both are cut into 2^25-word chunks (4 forward transforms, 3
    diagonal sums, 3 inverse transforms); also the square, whose B chunks are A's."""
    la = lb = 58_000_000
    assert engine.transform_log2(la, lb) == -1
    a = L.random_operand(28 * la - 5, 31)
    b = L.random_operand(28 * lb - 9, 32)
    p = engine.mul(a, b)
    assert p.size in (la + lb, la + lb - 1)
    check_residues3(a, b, p)
    check_residues3(a, None, engine.mul(a, None))


def test_beyond_one_transform_long_times_short():
    """A 2.24-Gbit operand (70 M words, beyond one transform) times a 28-Mbit one: chunk plans
    on 2^21-point transforms with one shared transform of the short operand."""
    la, lb = 80_000_000, 1_000_000
    assert engine.transform_log2(la, lb) == -1
    a = L.random_operand(28 * la, 41)
    b = L.random_operand(28 * lb, 42)
    p = engine.mul(a, b)
    check_residues3(a, b, p)


def _fast_operand(n: int, seed: int) -> np.ndarray:
    """n random 28-bit limbs with a nonzero top limb (numpy's generator: SplitMix64 repacking
    takes seconds per 10^8 limbs, and the residue checks do not depend on the generator)."""
    x = np.random.default_rng(seed).integers(0, 1 << 28, size=n, dtype=np.uint32)
    x[-1] |= 1 << 27
    return x


def test_beyond_three_primes_in_pieces_two():
    """Both operands above 2^26 words (76,695,851 limbs, 2.15 Gbit): one set of diagonal sums
    would exceed P0 P1 P2, so the shorter operand runs in two pieces (76,695,840 limbs and 11)
    whose products are added on the device; also the square, whose pieces are cut from A."""
    n = 76_695_851
    assert engine.transform_log2(n, n) == -1
    assert not engine.lib().bi_hip_plan_allowed(n, n, 0, 8)
    a = _fast_operand(n, 61)
    b = _fast_operand(n, 62)
    p = engine.mul(a, b)
    assert p.size in (2 * n, 2 * n - 1)
    check_residues3(a, b, p)
    check_residues3(a, None, engine.mul(a, None))


def test_beyond_three_primes_in_pieces_three():
    """4.3 Gbit x 4.4 Gbit: the shorter operand (153,396,680 limbs, after the operands are
    swapped) is cut into three pieces, the last of 5,000 limbs; two device additions carry the
    partial products into the output."""
    la, lb = 2 * 76_695_840 + 5_000, 2 * 76_695_840 + 3_000_000
    a = _fast_operand(la, 63)
    b = _fast_operand(lb, 64)
    p = engine.mul(a, b)
    assert p.size in (la + lb, la + lb - 1)
    check_residues3(a, b, p)


def test_pieces_product_inside_a_batch():
    """A pieces product in a grouped batch (run one by one after the batch's launch sets) next to
    small products, in table and learned modes: every product exact."""
    n = 76_695_851 + 700
    big_a, big_b = _fast_operand(n, 65), _fast_operand(n - 100, 66)
    rng = random.Random(67)
    pairs = [(L.random_operand(rng.randrange(3000, 90000), rng.getrandbits(63)),
              L.random_operand(rng.randrange(3000, 90000), rng.getrandbits(63))) for _ in range(6)]
    pairs.insert(3, (big_a, big_b))
    for mode in (engine.PLAN_TABLE, engine.PLAN_LEARNED):
        got = engine.mul_batch(pairs, mode)
        for i, ((a, b), p) in enumerate(zip(pairs, got)):
            if i == 3:
                check_residues3(a, b, p)
            else:
                assert np.array_equal(p, oracle.mul_limbs(a, b)), (mode, i)


def test_batch_graph_replays_match_eager():
    """Grouped batches run as captured graphs (the executable graph is updated in place when the
    next batch has the same shape).  The same batch three times as a graph, with eager batches
    and synchronizations in between, must give the eager products every time."""
    rng = random.Random(77)
    pairs = []
    for i in range(40):
        a = L.random_operand(int(2 ** rng.uniform(8, 17)), rng.getrandbits(63))
        b = None if i % 9 == 0 else L.random_operand(int(2 ** rng.uniform(8, 17)), rng.getrandbits(63))
        pairs.append((a, b))
    lib = engine.lib()
    lib.bi_hip_set_graphs(0)
    try:
        eager = engine.mul_batch(pairs, engine.PLAN_TABLE)
    finally:
        lib.bi_hip_set_graphs(1)
    for rep in range(3):
        got = engine.mul_batch(pairs, engine.PLAN_TABLE)
        for (a, b), e, g in zip(pairs, eager, got):
            assert np.array_equal(e, g), rep
        lib.bi_hip_set_graphs(0)
        try:
            other = engine.mul_batch(pairs[::-1], engine.PLAN_TABLE)  # eager work between replays
        finally:
            lib.bi_hip_set_graphs(1)
        assert all(np.array_equal(x, y) for x, y in zip(other, eager[::-1]))
    for (a, b), e in zip(pairs[:10], eager[:10]):
        assert np.array_equal(e, oracle.mul_limbs(a, b))


def test_profile_kinds_brackets_only_chosen_classes():
    """bi_hip_profile_kinds records events only around launches of the chosen kernel classes;
    the product is the same with every launch bracketed, with one class, and unprofiled."""
    a = L.random_operand(1 << 21, L.SEED)
    b = L.random_operand(1 << 21, L.SEED + 1)
    ref = engine.mul(a, b)
    engine.profile(True)
    p_all = engine.mul(a, b)
    every = engine.profile_read()
    engine.profile(False)
    used = [k for k, v in every.items() if v["launches"]]
    assert len(used) >= 2, every
    dom = max(used, key=lambda k: every[k]["ms"])
    engine.profile_kinds([dom])
    p_one = engine.mul(a, b)
    one = engine.profile_read()
    engine.profile_kinds([])
    assert one[dom]["launches"] == every[dom]["launches"]
    assert one[dom]["bytes"] == every[dom]["bytes"]
    assert all(v["launches"] == 0 for k, v in one.items() if k != dom), one
    assert np.array_equal(p_all, ref) and np.array_equal(p_one, ref)
    check_residues(a, b, ref)
