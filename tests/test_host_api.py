"""The host C++ API (include/bigint/bigint.h) driven through tests/cpp/bigint_calc.cpp.

CPU tests keep every product at or below 16,384 total bits, which the library multiplies on the
host exactly as the reference's dispatcher does (include/bigint/mul.h:80, 210-212); the GPU tests
push products through the engine.  Expected values come from CPython ints and from
tests/bigfloat_model.py, an independent Python model of the reference's BigFloat bookkeeping.
"""
import os
import random
import subprocess

import pytest

import bigfloat_model as BF

if hasattr(__import__("sys"), "set_int_max_str_digits"):
    __import__("sys").set_int_max_str_digits(0)  # decimal round trips of 60,000-digit numbers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALC = os.path.join(ROOT, "bigint_amd", "_build", "bigint_calc")


def hx(x: int) -> str:
    return ("-" if x < 0 else "") + format(abs(x), "x")


def unhx(s: str) -> int:
    return -int(s[1:], 16) if s.startswith("-") else int(s, 16)


class Calc:
    def __init__(self):
        self.p = subprocess.Popen([CALC], stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True,
                                  bufsize=1)

    def __call__(self, *args) -> str:
        self.p.stdin.write(" ".join(str(a) for a in args) + "\n")
        self.p.stdin.flush()
        return self.p.stdout.readline().strip()

    def close(self):
        self.p.stdin.close()
        self.p.wait(timeout=30)


@pytest.fixture(scope="module")
def calc():
    if not os.path.exists(CALC):
        pytest.skip("bigint_calc not built")
    c = Calc()
    yield c
    c.close()


def rnd(rng, maxbits):
    x = rng.getrandbits(rng.randrange(1, maxbits))
    return -x if rng.random() < 0.3 else x


def xf(f):
    neg, D, pp = f
    return f"{'-' if neg and D else '+'} {format(D, 'x')} {pp if D else 0}"


def trunc_div(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def divmod_mode(a, b, mode):
    if mode == 0:
        q = trunc_div(a, b)
    elif mode == 1:
        q = a // b
    elif mode == 2:
        q = -((-a) // b)
    else:  # half away from zero
        q0 = trunc_div(a, b)
        r = a - q0 * b
        q = q0 + ((1 if (a >= 0) == (b >= 0) else -1) if 2 * abs(r) >= abs(b) else 0)
    return q, a - q * b


def test_integer_ops_small(calc):
    rng = random.Random(1)
    for _ in range(300):
        a, b = rnd(rng, 4000), rnd(rng, 4000)
        assert unhx(calc("mul", hx(a), hx(b))) == a * b
        assert unhx(calc("add", hx(a), hx(b))) == a + b
        assert unhx(calc("sub", hx(a), hx(b))) == a - b
        assert unhx(calc("sq", hx(a))) == a * a
        k = rng.randrange(-(1 << 63), 1 << 63) or 3
        assert unhx(calc("muli", hx(a), k)) == a * k
        assert unhx(calc("divi", hx(a), k)) == trunc_div(a, k)
        s = rng.randrange(0, 300)
        sign = -1 if a < 0 else 1
        assert unhx(calc("shl", hx(a), s)) == sign * (abs(a) << s)
        assert unhx(calc("shr", hx(a), s)) == sign * (abs(a) >> s)
        assert int(calc("cmp", hx(a), hx(b))) == (a > b) - (a < b)
        assert unhx(calc("inc", hx(a))) == a + 1
        assert unhx(calc("dec", hx(a))) == a - 1
    assert unhx(calc("dec", "0")) == -1
    assert unhx(calc("inc", "-1")) == 0


def test_host_schoolbook_extremes(calc):
    """Products on the host route (<= 16,384 bits in total) with every limb all ones: the
    column sums of the schoolbook are at their largest (up to 586 terms of (2^28 - 1)^2)."""
    rng = random.Random(9)
    def words64(bits):
        return (bits + 63) // 64

    cases = [(8192, 8192), (8191, 8192), (16000, 300), (16320, 1), (28 * 292, 8192), (1, 1)]
    cases += [(rng.randrange(1, 8192), rng.randrange(1, 8192)) for _ in range(20)]
    for a_bits, b_bits in cases:
        assert (words64(a_bits) + words64(b_bits)) * 64 <= 16384  # the host route
        a, b = (1 << a_bits) - 1, (1 << b_bits) - 1
        assert unhx(calc("mul", hx(a), hx(b))) == a * b, (a_bits, b_bits)
        if 2 * words64(a_bits) * 64 <= 16384:
            assert unhx(calc("sq", hx(a))) == a * a, a_bits
    for _ in range(20):
        a = rng.getrandbits(8192) | (1 << 8191)
        b = rng.getrandbits(8192) | (1 << 8191)
        assert unhx(calc("mul", hx(a), hx(b))) == a * b


def test_bitwise_small(calc):
    rng = random.Random(2)
    for _ in range(100):
        a, b = rnd(rng, 500), rnd(rng, 500)
        sa = -1 if a < 0 else 1
        r = unhx(calc("and", hx(a), hx(b)))
        assert r == (sa * (abs(a) & abs(b)) if a else 0)
        r = unhx(calc("or", hx(a), hx(b)))
        assert r == (b if a == 0 else sa * (abs(a) | abs(b)))
        r = unhx(calc("xor", hx(a), hx(b)))
        assert r == (b if a == 0 else a if b == 0 else sa * (abs(a) ^ abs(b)))


def test_divmod_modes_small(calc):
    rng = random.Random(3)
    for _ in range(150):
        a, b = rnd(rng, 3000), rnd(rng, 1500)
        if b == 0:
            continue
        for mode in range(4):
            q, r = calc("divmod", hx(a), hx(b), mode).split()
            assert (unhx(q), unhx(r)) == divmod_mode(a, b, mode), (a, b, mode)


def test_decimal_round_trip_small(calc):
    rng = random.Random(4)
    for _ in range(50):
        a = rnd(rng, 3000)
        assert calc("todec", hx(a)) == str(a)
        assert unhx(calc("fromdec", str(a))) == a


def test_decimal_split_tree_small(calc):
    """Decimal output through the split tree (x = q 10^k + r, one depth at a time, quotients
    corrected until exact) on the host route: the brute-force threshold is lowered to a few limbs
    so that numbers of a few thousand bits split many levels deep, with zero runs to pad."""
    rng = random.Random(14)
    try:
        for thr in [2, 3, 5, 40]:
            assert calc("decthr", thr) == "ok"
            for _ in range(25):
                a = rnd(rng, 3000)
                assert calc("todec", hx(a)) == str(a), (thr, a)
            for z in [10 ** 600, 10 ** 600 - 1, 10 ** 500 + 7, 123 * 10 ** 300 + 10 ** 150 + 9,
                      -(10 ** 400 + 1), (1 << 2800) - 1]:
                assert calc("todec", hx(z)) == str(z), (thr, z)
    finally:
        assert calc("decthr", 2000) == "ok"


def test_bigfloat_mul_add_small(calc):
    rng = random.Random(5)
    for _ in range(200):
        a, b = rnd(rng, 2000), rnd(rng, 2000)
        pa, pb = rng.randrange(-5, 40), rng.randrange(-5, 40)
        fa, fb = BF.from_int(a, -28 * pa), BF.from_int(b, -28 * pb)
        prec = rng.choice([0, 1, 2, 5, 20, 70])
        assert calc("fmul", hx(a), pa, hx(b), pb, prec) == xf(BF.mul(fa, fb, prec))
        assert calc("fadd", hx(a), pa, hx(b), pb) == xf(BF.add_or_sub(fa, fb, False))
        assert calc("fsub", hx(a), pa, hx(b), pb) == xf(BF.add_or_sub(fa, fb, True))


def test_bigfloat_double_round_trip(calc):
    rng = random.Random(6)
    import struct
    for _ in range(200):
        x = rng.uniform(-1e30, 1e30) * 10 ** rng.randrange(-200, 200)
        bits = struct.unpack("<Q", struct.pack("<d", x))[0]
        f = BF.from_double(x)
        assert calc("ffromdouble", format(bits, "x")) == xf(f)
        neg, D, pp = f
        got = int(calc("ftodouble", hx(-D if neg else D), pp), 16)
        assert got == bits


def test_reciprocal_small(calc):
    """Limb-exact against the Python model, and within the Newton error bound."""
    rng = random.Random(7)
    for _ in range(40):
        a = rng.getrandbits(rng.randrange(30, 1500)) | 1
        pa = rng.randrange(-3, 30)
        prec = rng.choice([1, 2, 3, 8, 17, 40])
        f = BF.from_int(a, -28 * pa)
        exp = BF.reciprocal(f, prec)
        assert calc("frecip", hx(a), pa, prec) == xf(exp)
        # |x * b - 1| small relative to the working precision
        neg, D, pp = exp
        err = abs(D * a - (1 << (28 * (pp + pa))))
        assert err * (1 << (28 * max(1, prec - 2))) < (1 << (28 * (pp + pa)))


def test_print_fraction_small(calc):
    from fractions import Fraction
    rng = random.Random(8)
    for _ in range(40):
        a = rng.getrandbits(rng.randrange(10, 600))
        pa = rng.randrange(0, 12)
        digits = rng.randrange(1, 60)
        val = Fraction(a, 1 << (28 * pa))
        s = calc("fprint", hx(a), pa, digits)
        ip, fp = s.split(".")
        # round half up at `digits` decimals
        scaled = val * 10 ** digits
        q = int(scaled)
        if scaled - q >= Fraction(1, 2):
            q += 1
        assert int(ip) * 10 ** digits + int(fp) == q
        assert len(fp) == digits


def test_products_stay_on_host_below_16384_bits(calc):
    g0, h0 = map(int, calc("counts").split())
    calc("mul", "f" * 100, "e" * 100)
    g1, h1 = map(int, calc("counts").split())
    assert (g1, h1) == (g0, h0 + 1)
    # the largest host products, all-ones (every column of the schoolbook at its maximum)
    for bits_a, bits_b in [(8192, 8192), (16000, 320), (64, 16300)]:
        x, y = (1 << bits_a) - 1, (1 << bits_b) - 1
        assert unhx(calc("mul", hx(x), hx(y))) == x * y
    g2, _ = map(int, calc("counts").split())
    assert g2 == g1
    # single products stay on the host up to the measured crossover, 49,152 total bits (in 64-bit
    # words as the reference counts them); bigger ones go to the engine
    for bits_a, bits_b in [(24576, 24576), (40000, 9152)]:
        x, y = (1 << bits_a) - 1, (1 << bits_b) - 1
        assert unhx(calc("mul", hx(x), hx(y))) == x * y
    g3, _ = map(int, calc("counts").split())
    assert g3 == g1


@pytest.mark.gpu
def test_large_ops_through_engine(calc):
    rng = random.Random(9)
    g0, _ = map(int, calc("counts").split())
    for bits in [20000, 100000, 1 << 20]:
        a, b = rng.getrandbits(bits) | 1, -(rng.getrandbits(bits - 7) | 1)
        assert unhx(calc("mul", hx(a), hx(b))) == a * b
        assert unhx(calc("sq", hx(b))) == b * b
        for mode in range(4):
            q, r = calc("divmod", hx(a * 3 + 17), hx(b), mode).split()
            assert (unhx(q), unhx(r)) == divmod_mode(a * 3 + 17, b, mode)
    g1, _ = map(int, calc("counts").split())
    assert g1 > g0  # the big products went through the GPU engine
    a = rng.getrandbits(200000) | 1
    assert calc("todec", hx(a)) == str(a)
    assert calc("todec", hx(-a)) == str(-a)
    assert unhx(calc("fromdec", str(a))) == a
    # many pieces, with runs of zero digits that need padding inside the split
    z = 10 ** 70000 + 12345 * 10 ** 30000 + 7
    assert calc("todec", hx(z)) == str(z)
    assert calc("todec", hx(-z)) == str(-z)


@pytest.mark.gpu
def test_device_int_mul_add(calc):
    """DeviceInt (numbers kept in HBM between batches): A*B + C*D and A*A through the engine's
    device product batch and device additions, against Python ints; signs, carries into a new
    limb, exact cancellation to zero, and operands big enough for chunked transforms."""
    rng = random.Random(17)
    cases = []
    for bits in [30000, 200000, 3000000]:
        for _ in range(3):
            a, b = rnd(rng, bits) or 1, rnd(rng, bits) or 1
            c, d = rnd(rng, bits) or 1, rnd(rng, bits) or 1
            cases.append((a, b, c, d))
    x = rng.getrandbits(100000) | 1
    y = rng.getrandbits(90000) | 1
    cases.append((x, y, -x, y))                   # cancels to zero
    cases.append((x, y, x - 1, -y))               # y
    ones = (1 << 60000) - 1
    cases.append((ones, ones, ones, 2 * ones + 1))  # carries into a new limb
    for a, b, c, d in cases:
        s, sq = calc("devmuladd", hx(a), hx(b), hx(c), hx(d)).split()
        assert unhx(s) == a * b + c * d
        assert unhx(sq) == a * a


@pytest.mark.gpu
def test_large_reciprocal_matches_model(calc):
    rng = random.Random(10)
    for prec in [600, 3000]:
        a = rng.getrandbits(28 * 700) | 1
        f = BF.from_int(a, -28 * 650)
        assert calc("frecip", hx(a), 650, prec) == xf(BF.reciprocal(f, prec))


def _calc_session(lines):
    """Run bigint_calc in a fresh process (the thread pool is process state) over the lines."""
    if not os.path.exists(CALC):
        pytest.skip("bigint_calc not built")
    out = subprocess.run([CALC], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    return out.stdout.strip().split("\n")


def test_init_thread_pool_semantics():
    """init_thread_pool keeps the reference's semantics (bigint.h:35-38; ntt_multithread.cpp:
    287-290, 407-409): the first call wins, 0 and 1 both mean no worker threads, and a product in
    the reference's NTT band (more than 10,485,760 total bits) made before any call creates the pool
    at its default size, after which init_thread_pool has no effect.  A smaller product does not.
    (Without a GPU the large products here fail after the pool is fixed; the pool is the point.)"""
    hw = os.cpu_count()
    assert _calc_session(["poolinit 3", "poolinit 5", "pool"]) == ["ok", "ok", "3"]
    assert _calc_session(["poolinit 0", "pool"]) == ["ok", "1"]
    assert _calc_session(["poolinit 1", "poolinit 6", "pool"]) == ["ok", "ok", "1"]
    assert _calc_session(["pool", "poolinit 2", "pool"]) == [str(hw), "ok", str(hw)]
    big = hx((1 << 6_000_000) - 12345)        # 2 x 6 Mbit: the reference's NTT band
    mid = hx((1 << 3_000_000) - 777)          # 2 x 3 Mbit: its FFT band
    r = _calc_session([f"mul {big} {big}", "poolinit 3", "pool"])
    assert r[1:] == ["ok", str(hw)], r[1:]
    r = _calc_session([f"mul {mid} {mid}", "poolinit 3", "pool"])
    assert r[1:] == ["ok", "3"], r[1:]


def _calc_env_session(lines, env):
    if not os.path.exists(CALC):
        pytest.skip("bigint_calc not built")
    e = dict(os.environ)
    e.update(env)
    out = subprocess.run([CALC], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         timeout=300, env=e)
    assert out.returncode == 0, out.stderr
    return out.stdout.strip().split("\n")


def test_engine_failures_map_to_exceptions_without_gpu():
    """Engine failures reach the caller as exceptions, as the reference's do (bad_alloc from its
    allocations, bigint.h:115): without a GPU an engine-sized product raises
    bigint::mul::engine_error with the engine's status text (no CPU fallback), while host-sized
    products still work in the same process."""
    import shutil
    if shutil.which("rocm-smi") and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be present: the no-device path is not reachable")
    big = hx((1 << 40000) - 3)  # 80,000 total bits: above the host route's limit
    r = _calc_env_session([f"mul {big} {big}", "mul 7 6", f"fmul {big} 0 {big} 0 2000"], {})
    assert r[0].startswith("EXC ") and "bad_alloc" not in r[0], r[0]
    assert r[1] == "2a"
    assert r[2].startswith("EXC "), r[2]


@pytest.mark.gpu
def test_engine_out_of_memory_maps_to_bad_alloc():
    """With device allocations forced to fail from the host (BI_HIP_TEST_FAIL_ALLOC, read when the
    engine initialises), every engine route of the host library raises std::bad_alloc: a product,
    a precision-truncated product, a batch and a DeviceInt upload; host-sized products are
    unaffected.  The same commands succeed without the flag."""
    rng = random.Random(115)
    a, b = rng.getrandbits(200000) | 1, rng.getrandbits(150000) | 1
    lines = [f"mul {hx(a)} {hx(b)}", "mul 7 6", f"fmul {hx(a)} 3 {hx(b)} 5 1000",
             f"devmuladd {hx(a)} {hx(b)} {hx(b)} {hx(a)}", "batch 40 3 400000"]
    r = _calc_env_session(lines, {"BI_HIP_TEST_FAIL_ALLOC": "1"})
    assert r[0] == "EXC std::bad_alloc", r[0]
    assert r[1] == "2a"
    assert r[2] == "EXC std::bad_alloc", r[2]
    assert r[3] == "EXC std::bad_alloc", r[3]
    assert r[4] == "EXC std::bad_alloc", r[4]
    ok = _calc_env_session(lines, {})
    assert unhx(ok[0]) == a * b and ok[1] == "2a" and not any(x.startswith("EXC") for x in ok), ok


def _newton_bound_ok(a, pa, res, prec):
    """|x * b - 1| < 2^(-28 (prec - 2)) for x = D 2^(-28 pp) and b = a 2^(-28 pa): the error bound
    of the doubling Newton iteration (src/bigint.cpp:1044-1071) at its final precision."""
    neg, D, pp = res
    e = 28 * (pp + pa)
    err = abs(D * a - (1 << e)) if e >= 0 else abs(D * a * (1 << -e) - 1)
    return err * (1 << (28 * max(1, prec - 2))) < (1 << max(e, 0))


def _parse_f(s):
    sign, d, pp = s.split()
    return (sign == "-", int(d, 16), int(pp))


@pytest.mark.gpu
def test_reciprocal_at_engine_sizes(calc):
    """BigFloat::reciprocal at precisions of 2^12 limbs and more, where the Newton products run on
    the engine.  Pinned by what the mathematics fixes: the Newton error bound for random
    divisors; 1/(B-1) = sum B^-k (every limb 1) and 1/(B+1) = sum (B-1) B^-2k (limbs alternating
    B-1, 0), B = 2^28, up to the last few limbs; and 2^(28k-j) for b = 2^(j-28k), one nonzero
    limb.  Random divisors are also compared limb for limb with the Python model."""
    B = 1 << 28
    rng = random.Random(1044)
    g0, _ = map(int, calc("counts").split())
    for prec in [4096, 20000, 60000]:
        for a, pa in [(B - 1, 0), (B + 1, 0), (B - 1, 7), (B + 1, -3)]:
            res = _parse_f(calc("frecip", hx(a), pa, prec))
            assert _newton_bound_ok(a, pa, res, prec), (a, pa, prec)
            neg, D, pp = res
            n = (D.bit_length() + 27) // 28
            assert n >= prec - 2
            limbs = [(D >> (28 * i)) & (B - 1) for i in range(n)]
            top = limbs[4:]  # all but the lowest four limbs
            if a == B - 1:
                assert all(x == 1 for x in top), (pa, prec)
            else:
                pattern = [B - 1 if (n - 1 - i) % 2 == 0 else 0 for i in range(4, n)]
                assert top == pattern, (pa, prec)
        for k, j in [(0, 0), (1, 5), (300, 27), (5000, 13)]:
            res = _parse_f(calc("frecip", hx(1 << j), k, prec))  # b = 2^j * 2^(-28k)
            neg, D, pp = res
            # one nonzero limb (the final shift may leave a zero limb below it) and the value
            # exactly 1/b = 2^(28k - j)
            assert not neg and D & (D - 1) == 0 and D < (B << 28), (k, j, prec)
            assert D.bit_length() - 1 - 28 * pp == 28 * k - j, (k, j, prec)
    for prec, bits in [(4096, 28 * 5000), (20000, 28 * 12000), (30000, 28 * 30000)]:
        a = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        pa = rng.randrange(-50, 50)
        s = calc("frecip", hx(a), pa, prec)
        assert _newton_bound_ok(a, pa, _parse_f(s), prec), prec
        assert s == xf(BF.reciprocal(BF.from_int(a, -28 * pa), prec)), prec
    g1, _ = map(int, calc("counts").split())
    assert g1 > g0  # the Newton products went through the engine


@pytest.mark.gpu
def test_precision_products_at_engine_sizes(calc):
    """BigFloat::mul with 0 < precision < la + lb on operands of 20k-120k limbs, the engine route
    that fetches only the kept top limbs (include/bigint/mul.h:227-243, 302-312).  Cases: the cut
    inside a run of zero limbs of the product (stripped as trailing zeros), operands truncated to
    precision limbs whose new low limbs are zero, squares, and cuts at la + lb - 1.  Expected: the
    Python model of mul_digits (the exact product of the truncated operands, truncated again)."""
    B = 1 << 28
    rng = random.Random(302)
    g0, _ = map(int, calc("counts").split())
    X = rng.getrandbits(28 * 20000) | (1 << (28 * 20000 - 1))
    Y = rng.getrandbits(28 * 2000) | (1 << (28 * 2000 - 1)) | 1
    m = 3000
    a = X * B ** m + 1          # limbs 1 .. m-1 are zero
    la, lb = 20000 + m, 2000
    cases = [(a, Y, la), (a, Y, la + 1), (a, Y, la + lb - 1), (a, Y, la - 10), (a, Y, 20000 + 5)]
    P = rng.getrandbits(28 * 60000) | (1 << (28 * 60000 - 1))
    Q = rng.getrandbits(28 * 50000) | (1 << (28 * 50000 - 1))
    cases += [(P, Q, 30000), (P, Q, 80000), (P, Q, 109999), (P, -Q, 55000)]
    R = rng.getrandbits(28 * 120000) | (1 << (28 * 120000 - 1))
    S = rng.getrandbits(28 * 120000) | (1 << (28 * 120000 - 1))
    cases += [(R, S, 100000), (R, S, 239999)]
    for x, y, prec in cases:
        px, py = rng.randrange(-40, 40), rng.randrange(-40, 40)
        fx, fy = BF.from_int(x, -28 * px), BF.from_int(y, -28 * py)
        assert calc("fmul", hx(x), px, hx(y), py, prec) == xf(BF.mul(fx, fy, prec)), prec
    for x, prec in [(a, la), (a, la + 7), (P, 45000), (R, 200000)]:
        px = rng.randrange(-40, 40)
        fx = BF.from_int(x, -28 * px)
        assert calc("fsqr", hx(x), px, prec) == xf(BF.mul(fx, fx, prec, same=True)), prec
    g1, _ = map(int, calc("counts").split())
    assert g1 > g0


def _splitmix(seed):
    st = seed
    M = (1 << 64) - 1
    while True:
        st = (st + 0x9e3779b97f4a7c15) & M
        z = st
        z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & M
        z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & M
        yield z ^ (z >> 31)


def _batch_checksum(n, seed, maxbits):
    g = _splitmix(seed)

    def operand():
        bits = 1 + next(g) % maxbits
        v = 0
        for w in range((bits + 63) // 64):
            x = next(g)
            top = bits - 64 * w
            if top <= 64:
                if top < 64:
                    x &= (1 << top) - 1
                x |= 1 << (top - 1)
            v |= x << (64 * w)
        return v

    M = (1 << 61) - 1
    s = 0
    for i in range(n):
        a = operand()
        b = a if i % 7 == 0 else operand()
        s = (s + (a * b % M) * (i + 1)) % M
    return s


def test_batch_of_host_products_on_pool_threads(calc):
    """BigInt::mul_batch with 20,000 host-sized pairs (< 16,384 total bits): the products run on
    the init_thread_pool threads; each slot must hold exactly its own product."""
    g0, h0 = map(int, calc("counts").split())
    assert int(calc("batch", 20000, 77, 8000)) == _batch_checksum(20000, 77, 8000)
    g1, h1 = map(int, calc("counts").split())
    assert g1 == g0 and h1 - h0 == 20000


def test_limb_passes_on_worker_threads(calc):
    """Additions, subtractions, small-factor products and shifts of operands above a million
    limbs run in chunks on worker threads, carries settled between chunks afterwards: compare them
    with CPython ints and the BigFloat model, including carries and borrows that ripple through
    every chunk (all-ones and power-of-two operands)."""
    rng = random.Random(11)
    n = 2_600_000  # limbs
    big = rng.getrandbits(28 * n) | (1 << (28 * n - 1))
    big2 = rng.getrandbits(28 * n - 5) | 1
    ones = (1 << (28 * n)) - 1
    power = 1 << (28 * n)
    # BigInt magnitude passes
    for a, b in ((big, big2), (ones, 1), (power, -1), (big, -big2), (-big2, big)):
        assert unhx(calc("add", hx(a), hx(b))) == a + b
        assert unhx(calc("sub", hx(a), hx(b))) == a - b
    for a, k in ((big, 10005), (ones, 3), (big2, (1 << 62) + 12345), (-ones, -(1 << 40) - 7)):
        assert unhx(calc("muli", hx(a), k)) == a * k
    # BigFloat passes (frames offset by a few limbs, so the operands overlap partially)
    for a, pa, b, pb in ((big, 3, big2, 7), (ones, 0, 1, 0), (power, 2, 1, 2), (-big, 5, big2, 1)):
        fa, fb = BF.from_int(a, -28 * pa), BF.from_int(b, -28 * pb)
        assert calc("fadd", hx(a), pa, hx(b), pb) == xf(BF.add_or_sub(fa, fb, False))
        assert calc("fsub", hx(a), pa, hx(b), pb) == xf(BF.add_or_sub(fa, fb, True))
    for a, pa, k in ((big, 4, 10005), (ones, 1, 7), (-big2, 2, (1 << 50) + 3)):
        neg, D, pp = BF.from_int(a, -28 * pa)
        assert calc("fmuli", hx(a), pa, k) == xf((neg != (k < 0), D * abs(k), pp))
    # (shift_right runs where the low limb has enough trailing zero bits: big << 20 by 13)
    for a, pa, k in ((big, 4, -13), (big << 20, 4, -13), (big, 4, 13), (ones, 0, -1), (big2 | (1 << 200), 3, 27)):
        assert calc("fshift", hx(a), pa, k) == xf(BF.shift(BF.from_int(a, -28 * pa), k))
