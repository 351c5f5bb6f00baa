"""Independent Python model of the reference's BigFloat bookkeeping (src/bigint.cpp), used to pin
the host library's BigFloat results limb for limb.

A BigFloat is (neg, D, pp): magnitude limbs = the 28-bit limbs of the Python int D (leading
zeros never stored; low zero limbs are part of D), value = (-1)^neg * D * 2^(-28 pp).
"""
from __future__ import annotations

import math
import struct

DB = 28
M = (1 << DB) - 1


def nlimbs(D: int) -> int:
    return max(1, -(-D.bit_length() // DB))


def limb(D: int, i: int) -> int:
    return (D >> (DB * i)) & M


def canon_tail(f):
    """remove_tail_zero (src/bigint.cpp:834-847)."""
    neg, D, pp = f
    if D == 0:
        return (False, 0, 0)
    z = ((D & -D).bit_length() - 1) // DB  # whole zero limbs at the bottom
    return (neg, D >> (DB * z), pp - z)


def from_int(x: int, offset: int = 0):
    """BigFloat(BigInt x, offset) = shift then remove_tail_zero."""
    f = (x < 0, abs(x), 0)
    f = shift(f, offset)
    return canon_tail(f)


def _trunc_div(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def _trunc_mod(a, b):
    return a - b * _trunc_div(a, b)


def shift(f, offset):
    """BigFloat::shift (src/bigint.cpp:769-812), C++ division semantics."""
    neg, D, pp = f
    pp -= _trunc_div(offset, DB)
    off = _trunc_mod(offset, DB)
    if off == 0:
        return (neg, D, pp) if D else (False, 0, 0)
    if off > 0:
        off = DB - off
        pp -= 1
    else:
        off = -off
    d0 = limb(D, 0)
    ctz = 32 if d0 == 0 else (d0 & -d0).bit_length() - 1
    if ctz >= off:
        D >>= off
    else:
        D <<= DB - off
        pp += 1
    if D == 0:
        return (False, 0, 0)
    return (neg, D, pp)


def to_double(f) -> float:
    """BigFloat::to_double: the mantissa is truncated, not rounded (src/bigint.cpp:717-766)."""
    neg, D, pp = f
    if D == 0:
        return 0.0
    e = D.bit_length() - 1 - DB * pp  # exponent of the top bit
    top = D >> max(0, D.bit_length() - 53)
    if D.bit_length() < 53:
        top = D << (53 - D.bit_length())
    ebits = e + 1023
    if ebits >= 2047:
        return -math.inf if neg else math.inf
    if ebits <= 0:
        raise NotImplementedError("subnormal not modelled")
    bits = (int(neg) << 63) | (ebits << 52) | (top & ((1 << 52) - 1))
    return struct.unpack("<d", struct.pack("<Q", bits))[0]


def from_double(x: float):
    """BigFloat(double) (src/bigint.cpp:681-715)."""
    if x == 0:
        return (False, 0, 0)
    bits = struct.unpack("<Q", struct.pack("<d", x))[0]
    neg = bool(bits >> 63)
    eb = (bits >> 52) & 0x7FF
    mant = bits & ((1 << 52) - 1)
    offset = -52 - 1023
    if eb == 0:
        offset += 1
    else:
        mant |= 1 << 52
        offset += eb
    return shift((neg, mant, 0), offset)


def add_or_sub(a, b, is_sub):
    """BigFloat::add_or_sub: exact, canonical (no leading/trailing zero limbs)."""
    an, aD, ap = a
    bn, bD, bp = b
    if aD == 0:
        return (not bn, bD, bp) if (is_sub and bD) else b
    if bD == 0:
        return a
    # common scale: lowest point position among nonzero low limbs
    ta = canon_tail(a)
    tb = canon_tail(b)
    pp = max(ta[2], tb[2])
    va = ta[1] << (DB * (pp - ta[2]))
    vb = tb[1] << (DB * (pp - tb[2]))
    sa = -va if an else va
    sb = -vb if bn else vb
    s = sa - sb if is_sub else sa + sb
    if s == 0:
        return (False, 0, 0)
    return canon_tail((s < 0, abs(s), pp))


def mul_digits_model(a: int, b: int, same: bool, precision: int):
    """mul.h:223-318 on limb ints: returns (product D, point delta)."""
    la, lb = nlimbs(a), nlimbs(b)
    oa = 0 if precision == 0 or la <= precision else la - precision
    ob = 0 if precision == 0 or lb <= precision else lb - precision
    za = 0
    while limb(a, oa + za) == 0:
        za += 1
    zb = za if same else 0
    if not same:
        while limb(b, ob + zb) == 0:
            zb += 1
    oa += za
    ob += zb
    res = (a >> (DB * oa)) * (b >> (DB * ob))
    if precision:
        lr = nlimbs(res)
        orr = 0 if lr <= precision else lr - precision
        while limb(res, orr) == 0:
            orr += 1
        return res >> (DB * orr), oa + ob + orr
    return res << (DB * (za + zb)), 0


def mul(a, b, precision, same=False):
    an, aD, ap = a
    bn, bD, bp = b
    if aD == 0 or bD == 0:
        return (False, 0, 0)
    D, delta = mul_digits_model(aD, bD, same, precision)
    return canon_tail((an != bn, D, ap + bp - delta))


def reciprocal(f, precision=0):
    """BigFloat::reciprocal (src/bigint.cpp:1044-1071)."""
    neg, D, pp = f
    if precision == 0:
        precision = nlimbs(D)
    exponent = DB * (nlimbs(D) - 1 - pp) + limb(D, nlimbs(D) - 1).bit_length()
    b = canon_tail(shift(f, -exponent))
    x = from_double(1 / to_double(b))
    cur = 52 // DB
    one = from_int(1)
    while cur < precision:
        cur *= 2
        x = add_or_sub(x, mul(x, add_or_sub(one, mul(b, x, cur + 1), True), cur + 1), False)
    return shift(x, -exponent)
