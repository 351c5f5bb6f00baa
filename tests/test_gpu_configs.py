"""The batched workload as it is benchmarked, and products queued on several streams.

Batched workload: 4096 independent products whose operand lengths are log-uniform in
[2^12, 2^24] bits (limbs.batch_bits(4096)), SplitMix64 operands (limbs.batch_seeds), run
through bi_hip_mul_batch_device in the plain (one launch sequence per product), autotuned-table
and learned plan modes.  Bar: the three modes give identical limbs, every product satisfies
(a mod q)(b mod q) == (a*b) mod q for six primes q below 2^31, and 64 products (the 8 largest,
the 8 smallest and 48 drawn at random) equal the oracle's limb for limb.
"""
import random

import numpy as np
import pytest

from bigint_amd import engine, hip, planner
from bigint_amd import limbs as L
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init(0)
    import os
    if os.path.exists(planner.DEFAULT_PATH):
        planner.load()


def _batch_pairs(pool_ptr, offs, lens, out_ptr, out_offs):
    return [(pool_ptr + 4 * int(offs[2 * i]), int(lens[2 * i]), pool_ptr + 4 * int(offs[2 * i + 1]),
             int(lens[2 * i + 1]), out_ptr + 4 * int(out_offs[i]), int(lens[2 * i] + lens[2 * i + 1]))
            for i in range(len(out_offs))]


def test_c3_batch_4096_all_modes():
    count = 4096
    bits = L.batch_bits(count)
    seeds = L.batch_seeds(count)
    flat_bits = [x for p in bits for x in p]
    flat_seeds = [x for p in seeds for x in p]
    pool, offs, lens = oracle.random_operands(flat_bits, flat_seeds)
    la, lb = lens[0::2], lens[1::2]
    out_lens = la + lb
    out_offs = np.concatenate([[0], np.cumsum(out_lens)[:-1]]).astype(np.uint64)
    total_out = int(out_lens.sum())
    print(f"C3: {count} pairs, {sum(flat_bits) / 1e9:.2f} Gbit of operands, {pool.size} limbs", flush=True)
    d_in = hip.Buffer.from_array(pool)
    stream = hip.Stream()
    outs = {}
    for mode in (engine.PLAN_FIXED, engine.PLAN_TABLE, engine.PLAN_LEARNED):
        d_out = hip.Buffer(4 * total_out)
        engine.mul_batch_device(_batch_pairs(d_in.ptr, offs, lens, d_out.ptr, out_offs), mode, stream.handle)
        stream.synchronize()
        outs[mode] = d_out.download(np.uint32, total_out)
        d_out.free()
        print(f"mode {mode} done", flush=True)
    d_in.free()
    assert np.array_equal(outs[engine.PLAN_FIXED], outs[engine.PLAN_TABLE])
    assert np.array_equal(outs[engine.PLAN_FIXED], outs[engine.PLAN_LEARNED])
    prod = outs[engine.PLAN_TABLE]
    for q in L.RESIDUE_PRIMES:
        r_in = oracle.residues(pool, offs, lens, q)
        r_out = oracle.residues(prod, out_offs, out_lens.astype(np.uint64), q)
        bad = np.flatnonzero(r_in[0::2] * r_in[1::2] % q != r_out)
        assert bad.size == 0, (q, bad[:10])
    # limb for limb against the oracle: the 8 largest and 8 smallest products, 48 more at random
    order = np.argsort(out_lens, kind="stable")
    pick = set(order[:8].tolist()) | set(order[-8:].tolist())
    rng = random.Random(2024)
    while len(pick) < 64:
        pick.add(rng.randrange(count))
    for i in sorted(pick):
        a = pool[int(offs[2 * i]):int(offs[2 * i] + la[i])]
        b = pool[int(offs[2 * i + 1]):int(offs[2 * i + 1] + lb[i])]
        got = prod[int(out_offs[i]):int(out_offs[i] + out_lens[i])]
        exp = oracle.mul_limbs(a, b)
        assert np.array_equal(L.strip(got), exp), (i, a.size, b.size)
        assert got.size == exp.size or (got.size == exp.size + 1 and got[-1] == 0)


def test_products_on_two_streams_without_host_sync():
    """Products queued on two streams (and the engine's own stream) back to back, with no host
    synchronisation between the calls: they share the engine's workspace, so each call has to
    wait on the device for the previous one.  Every product must equal the oracle's."""
    a1 = L.random_operand(1 << 22, 11)
    b1 = L.random_operand(1 << 22, 12)
    a2 = L.random_operand((1 << 21) + 77, 13)
    b2 = L.random_operand(1 << 20, 14)
    a3 = L.random_operand(28 * 60000, 15)   # chunked through the batch path (descriptor buffer)
    b3 = L.random_operand(28 * 2000, 16)
    ops = [(a1, b1), (a2, b2), (a3, b3)]
    exp = [oracle.mul_limbs(a, b) for a, b in ops]
    bufs = [(hip.Buffer.from_array(a), hip.Buffer.from_array(b), hip.Buffer(4 * (a.size + b.size)))
            for a, b in ops]
    sx, sy = hip.Stream(), hip.Stream()

    def queue(i, st):
        a, b = ops[i]
        da, db, do = bufs[i]
        if i == 2:
            engine.mul_batch_device([(da.ptr, a.size, db.ptr, b.size, do.ptr, a.size + b.size)],
                                    engine.PLAN_TABLE, st.handle)
        else:
            engine.mul_device(da.ptr, a.size, db.ptr, b.size, False, do.ptr, a.size + b.size, st.handle)

    # once each, synchronised: the workspace reaches its largest size here, so the calls below
    # do not grow it (growing waits for the whole device)
    for i in range(3):
        queue(i, sx)
        sx.synchronize()
    for rep in range(3):
        for d in (bufs[0][2], bufs[1][2], bufs[2][2]):
            d.zero()
        hip.synchronize()
        order = [(0, sx), (1, sy), (2, sx), (0, sy)] if rep % 2 == 0 else [(1, sx), (0, sy), (2, sy), (1, sx)]
        for i, st in order:
            queue(i, st)
        # a product through the host API on the engine's stream, between the queued ones
        host = engine.mul(a2, b2)
        sx.synchronize()
        sy.synchronize()
        assert np.array_equal(host, exp[1])
        for i, (a, b) in enumerate(ops):
            got = bufs[i][2].download(np.uint32, a.size + b.size)
            assert np.array_equal(L.strip(got), exp[i]), (rep, i)
    for t in bufs:
        for x in t:
            x.free()


def test_back_to_back_batches_without_host_sync():
    """Grouped batches of different shapes queued back to back, on one stream and on two streams
    in turn, with no host synchronisation between the calls: a batch call does not wait for the
    earlier ones, its descriptors reach the device with a copy ordered after the earlier batches
    that read the descriptor buffer, and the two executable graphs are used in turn (updated in
    place or instantiated anew as the shapes change).  Every product must equal the oracle's."""
    rng = np.random.default_rng(5)
    batches = []
    for bi in range(4):
        pairs = []
        for k in range(24 + 8 * bi):
            ba, bb = int(2 ** rng.uniform(12, 19)), int(2 ** rng.uniform(12, 19))
            pairs.append((L.random_operand(ba, 1000 * bi + 2 * k), L.random_operand(bb, 1000 * bi + 2 * k + 1)))
        batches.append(pairs)
    exp = [[oracle.mul_limbs(a, b) for a, b in pairs] for pairs in batches]
    bufs = [[(hip.Buffer.from_array(a), hip.Buffer.from_array(b), hip.Buffer(4 * (a.size + b.size)))
             for a, b in pairs] for pairs in batches]
    sx, sy = hip.Stream(), hip.Stream()

    def queue(i, st, mode):
        engine.mul_batch_device([(da.ptr, a.size, db.ptr, b.size, do.ptr, a.size + b.size)
                                 for (a, b), (da, db, do) in zip(batches[i], bufs[i])], mode, st.handle)

    # once each, synchronised: the workspace and descriptor buffers reach their largest size here
    for i in range(4):
        queue(i, sx, engine.PLAN_TABLE)
        sx.synchronize()
    orders = [[(0, sx), (1, sx), (2, sx), (3, sx), (0, sx), (2, sx)],
              [(1, sx), (3, sy), (0, sx), (2, sy), (1, sy), (3, sx)]]
    for rep, order in enumerate(orders):
        for t in bufs:
            for _, _, do in t:
                do.zero()
        hip.synchronize()
        for k, (i, st) in enumerate(order):
            queue(i, st, engine.PLAN_LEARNED if (k + rep) % 2 else engine.PLAN_TABLE)
        sx.synchronize()
        sy.synchronize()
        for i, pairs in enumerate(batches):
            for j, (a, b) in enumerate(pairs):
                got = bufs[i][j][2].download(np.uint32, a.size + b.size)
                assert np.array_equal(L.strip(got), exp[i][j]), (rep, i, j)
    for t in bufs:
        for x in t:
            for y in x:
                y.free()
