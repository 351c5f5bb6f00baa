"""Limb layout helpers and the seeded operand generator (CPU)."""
import numpy as np
import pytest

from bigint_amd import limbs as L


def test_splitmix64_known_answer():
    # published SplitMix64 outputs for seed 1234567
    assert [int(x) for x in L.splitmix64(1234567, 5)] == [
        6457827717110365317, 3203168211198807973, 9817491932198370423,
        4593380528125082431, 16408922859458223821]
    s, out = 1234567, []
    for _ in range(3):
        s, z = L.splitmix64_scalar(s)
        out.append(z)
    assert out == [int(x) for x in L.splitmix64(1234567, 3)]


@pytest.mark.parametrize("bits", [1, 27, 28, 29, 63, 64, 65, 447, 448, 449, 1000, 4096, 100003])
def test_operand_bits(bits):
    a = L.random_operand(bits, L.SEED + bits)
    x = L.limbs_to_int(a)
    assert x.bit_length() == bits
    assert a.size == -(-bits // 28)
    assert int(a.max()) < (1 << 28)
    assert np.array_equal(L.int_to_limbs(x), a)


def test_int_round_trip():
    rng = np.random.default_rng(1)
    for nb in [1, 2, 28, 56, 448, 1000, 12345]:
        x = int.from_bytes(rng.bytes(nb), "little") | 1
        assert L.limbs_to_int(L.int_to_limbs(x)) == x
    assert list(L.int_to_limbs(0)) == [0]
    assert list(L.int_to_limbs((1 << 28) - 1)) == [(1 << 28) - 1]
    assert list(L.int_to_limbs(1 << 28)) == [0, 1]


def test_words_limbs_round_trip():
    w = L.splitmix64(7, 33)
    limbs = L.words_to_limbs(w)
    back = L.limbs_to_words(limbs)
    assert np.array_equal(back[: w.size], w)
    assert not back[w.size:].any()


def test_residue_mod_matches_python():
    a = L.random_operand(50000, 11)
    x = L.limbs_to_int(a)
    for q in L.RESIDUE_PRIMES:
        assert L.residue_mod(a, q) == x % q


def test_batch_bits_range():
    bb = L.batch_bits(4096)
    flat = [b for p in bb for b in p]
    assert min(flat) >= 1 << 12 and max(flat) < 1 << 24
    # log-uniform: median near 2^18
    med = sorted(flat)[len(flat) // 2]
    assert (1 << 17) < med < (1 << 19)


def test_batch_checker_matches_python():
    """oracle/batch_check.c (seeded operands and residues for gigabit batches) against
    limbs.random_operand and limbs.residue_mod."""
    from oracle import oracle
    bits = [1, 27, 28, 29, 63, 64, 65, 1000, 123457, 1 << 20]
    seeds = [L.SEED + i for i in range(len(bits))]
    pool, offs, lens = oracle.random_operands(bits, seeds)
    for b, s, o, n in zip(bits, seeds, offs, lens):
        assert np.array_equal(pool[int(o):int(o + n)], L.random_operand(b, s)), b
    for q in L.RESIDUE_PRIMES:
        r = oracle.residues(pool, offs, lens, q, threads=3)
        for i, (o, n) in enumerate(zip(offs, lens)):
            assert r[i] == L.residue_mod(pool[int(o):int(o + n)], q)
    assert L.batch_seeds(2) == [(L.SEED + 2, L.SEED + 3), (L.SEED + 4, L.SEED + 5)]
