"""28-bit limb helpers and the seeded operand generator.

Limb layout is the reference's ``Digits``: ``std::vector<uint32_t>``, little-endian, each limb
below 2^28, value = sum(d[i] * 2^(28*i)) (include/bigint/bigint_base.h:21-26).

Operands for the benchmark workloads come from SplitMix64: operand A uses seed
0x9e3779b97f4a7c15 (the seed of the reference's regression, tests/measure_digit_bits.cpp:140),
B uses seed+1 and batch pair i uses seeds seed+2+2i and seed+3+2i (batch_seeds).  An operand of
b bits takes ceil(b/64) words of the stream, is masked to b bits and gets bit b-1 set, then is
repacked into 28-bit limbs.
"""
from __future__ import annotations

import numpy as np

DIGIT_BITS = 28
DIGIT_MASK = (1 << DIGIT_BITS) - 1
SEED = 0x9E3779B97F4A7C15
_GAMMA = 0x9E3779B97F4A7C15
_M64 = (1 << 64) - 1


def splitmix64(seed: int, count: int) -> np.ndarray:
    """The first ``count`` outputs of SplitMix64 started from ``seed`` (uint64 array)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & _M64) + idx * np.uint64(_GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def splitmix64_scalar(state: int) -> tuple[int, int]:
    """One SplitMix64 step on a Python int state: returns (new_state, output)."""
    state = (state + _GAMMA) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


def words_to_limbs(words: np.ndarray, nbits: int | None = None) -> np.ndarray:
    """Little-endian uint64 words -> 28-bit limbs (uint32), trimmed to ceil(nbits/28) limbs
    when ``nbits`` is given, else leading zero limbs are stripped (at least one limb kept)."""
    words = np.ascontiguousarray(words, dtype=np.uint64)
    total_bits = words.size * 64
    nl = -(-total_bits // DIGIT_BITS)
    pos = np.arange(nl, dtype=np.uint64) * np.uint64(DIGIT_BITS)
    wi = (pos >> np.uint64(6)).astype(np.int64)
    off = pos & np.uint64(63)
    padded = np.concatenate([words, np.zeros(2, dtype=np.uint64)])
    lo = padded[wi] >> off
    # bits that spill into the next word; shifting by 64 is undefined, so mask off == 0
    sh = (np.uint64(64) - off) & np.uint64(63)
    hi = np.where(off == 0, np.uint64(0), padded[wi + 1] << sh)
    limbs = ((lo | hi) & np.uint64(DIGIT_MASK)).astype(np.uint32)
    if nbits is not None:
        return limbs[: -(-nbits // DIGIT_BITS)].copy()
    return strip(limbs)


def limbs_to_words(limbs: np.ndarray) -> np.ndarray:
    """28-bit limbs -> little-endian uint64 words (16 limbs = 448 bits = 7 words per group)."""
    limbs = np.ascontiguousarray(limbs, dtype=np.uint32)
    g = -(-limbs.size // 16)
    d = np.zeros(g * 16, dtype=np.uint64)
    d[: limbs.size] = limbs
    d = d.reshape(g, 16)
    out = np.zeros((g, 7), dtype=np.uint64)
    for k in range(16):
        bit = 28 * k
        w, off = divmod(bit, 64)
        out[:, w] |= d[:, k] << np.uint64(off)
        if off + 28 > 64:
            out[:, w + 1] |= d[:, k] >> np.uint64(64 - off)
    return out.reshape(-1)


def strip(limbs: np.ndarray) -> np.ndarray:
    """Drop leading (most significant) zero limbs, keeping at least one limb."""
    nz = np.flatnonzero(limbs)
    n = int(nz[-1]) + 1 if nz.size else 1
    return limbs[:n]


def int_to_limbs(x: int) -> np.ndarray:
    """Non-negative Python int -> canonical 28-bit limbs (zero is [0])."""
    if x < 0:
        raise ValueError("magnitude must be non-negative")
    if x == 0:
        return np.zeros(1, dtype=np.uint32)
    nbytes = (x.bit_length() + 7) // 8
    nbytes8 = -(-nbytes // 8) * 8
    words = np.frombuffer(x.to_bytes(nbytes8, "little"), dtype="<u8").astype(np.uint64)
    return words_to_limbs(words, x.bit_length())


def limbs_to_int(limbs: np.ndarray) -> int:
    """28-bit limbs -> Python int."""
    words = limbs_to_words(np.asarray(limbs, dtype=np.uint32))
    return int.from_bytes(words.astype("<u8").tobytes(), "little")


def random_operand(bits: int, seed: int) -> np.ndarray:
    """Operand of exactly ``bits`` bits from SplitMix64(seed): ceil(bits/64) words, masked to
    ``bits`` bits, top bit set, as 28-bit limbs (ceil(bits/28) of them)."""
    if bits <= 0:
        raise ValueError("bits must be positive")
    nw = -(-bits // 64)
    w = splitmix64(seed, nw)
    r = bits % 64
    if r:
        w[-1] &= np.uint64((1 << r) - 1)
    w[-1] |= np.uint64(1 << ((bits - 1) % 64))
    return words_to_limbs(w, bits)


def batch_bits(count: int, seed: int = SEED, lo_exp: int = 12, span: int = 12) -> list[tuple[int, int]]:
    """Per-pair operand bit lengths for the batched workload: for each operand draw
    u = (x >> 11) * 2^-53 from SplitMix64(seed) and take floor(2^(lo_exp + span*u))."""
    xs = splitmix64(seed ^ 0x5BD1E995, 2 * count)
    u = (xs >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    bits = np.floor(np.exp2(lo_exp + span * u)).astype(np.int64)
    return [(int(bits[2 * i]), int(bits[2 * i + 1])) for i in range(count)]


def batch_seeds(count: int, seed: int = SEED) -> list[tuple[int, int]]:
    """SplitMix64 seeds of the batched workload's operands: pair i takes seed + 2 + 2i for its
    first operand and seed + 3 + 2i for its second (seed and seed + 1 are the single product's)."""
    return [((seed + 2 + 2 * i) & _M64, (seed + 3 + 2 * i) & _M64) for i in range(count)]


# Primes below 2^31 for size-independent product checks: (a mod q)(b mod q) = (a*b) mod q.
RESIDUE_PRIMES = (2147483629, 2147483587, 2147483579, 2147483563, 2147483549, 2147483543)


def residue_mod(limbs: np.ndarray, q: int) -> int:
    """sum(limb_i * 2^(28 i)) mod q for q < 2^31, vectorised (O(n)), in chunks of 2^22 limbs so
    that products of billions of bits need little host memory."""
    limbs = np.asarray(limbs)
    n = limbs.size
    if n == 0:
        return 0
    qq = np.uint64(q)
    chunk = min(n, 1 << 22)
    w = np.ones(1, dtype=np.uint64)
    step = (1 << 28) % q  # 2^(28*len(w)) mod q
    while w.size < chunk:
        w = np.concatenate([w, (w * np.uint64(step)) % qq])
        step = step * step % q
    w = w[:chunk]
    shift = pow(2, 28 * chunk, q)  # weight of a chunk relative to the one below it
    r = 0
    for start in range(((n - 1) // chunk) * chunk, -1, -chunk):  # top chunk first (Horner)
        part = limbs[start:start + chunk].astype(np.uint64)
        if q <= 1 << 28:  # a limb (below 2^28) may exceed q
            part %= qq
        terms = part * w[:part.size] % qq  # each < 2^31: 2^22 of them sum below 2^53
        r = (r * shift + int(terms.sum(dtype=np.uint64))) % q
    return r
