"""CPU model of the HIP engine's transform schedule (bigint_amd/csrc/ntt_kernels.hip,
bigint_amd/csrc/engine.cpp), run against CPython int products.

The model replays the kernels' index arithmetic literally (thread -> register-group layout,
LDS exchange addresses, per-tile twiddle tables tw[k] = W[k] * beta^tr, outer/middle pass
tiling, Garner pieces and the map-based carry scan) on small transform lengths, and asserts every
global and LDS index stays inside its buffer.  It checks the schedule's mathematics and bounds on
the CPU before any of it runs on a GPU; it does not execute the GPU code.
"""
import random

import pytest

from bigint_amd import limbs as L

P = [469762049, 2013265921, 2281701377]
G = [3, 31, 3]
R32 = 1 << 32
KE, KTILE = 16, 4096
KLDS = KTILE + KTILE // 16
MASK = (1 << 28) - 1


def mont(I, a, b):
    return a * b * pow(R32, -1, P[I]) % P[I]


def to_mont(I, x):
    return x * R32 % P[I]


def bitrev(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


def root(I, k, inv):
    w = pow(G[I], (P[I] - 1) >> k, P[I])
    return pow(w, -1, P[I]) if inv else w


def wtab(I, inv):
    t = [to_mont(I, 1)] * KTILE
    lm = 0
    while (1 << lm) < KTILE:
        m = 1 << lm
        w = root(I, lm + 1, inv)
        for i in range(m):
            t[m + i] = to_mont(I, pow(w, bitrev(i, lm), P[I]))
        lm += 1
    return t


def beta_table(I, log_n, log_s, nblk, inv):
    if nblk <= 1:
        return None
    lb = nblk.bit_length() - 1
    w = root(I, log_n - log_s, inv)
    return [to_mont(I, pow(w, bitrev(b, lb), P[I])) for b in range(nblk)]


def build_tw(I, W, beta, log_r):
    R = 1 << log_r
    if beta is None:
        return list(W[:R])
    bpow = []
    b = beta
    for e in range(log_r):
        bpow.append(b)
        b = mont(I, b, b)
    tw = [0] * R
    for k in range(1, R):
        log_m = k.bit_length() - 1
        tw[k] = mont(I, W[k], bpow[log_r - 1 - log_m])
    return tw


def groups(log_r):
    ng = (log_r + 3) // 4
    out = []
    for gi in range(ng):
        st = min(4, log_r - 4 * gi)
        out.append((st, log_r - 4 * gi - st))
    return out


def layout(EG, log_c, log_trlo, nthreads, tid):
    le = EG.bit_length() - 1
    res = []
    for sg in range(KE // EG):
        s = tid + nthreads * sg
        col = s & ((1 << log_c) - 1)
        rest = s >> log_c
        rl = rest & ((1 << log_trlo) - 1)
        rh = rest >> log_trlo
        res.append(((rh << (log_trlo + le)) + rl, col))
    return res


def rows_cols(EG, log_c, log_trlo, nthreads, tid):
    out = []
    for base, col in layout(EG, log_c, log_trlo, nthreads, tid):
        for q in range(EG):
            out.append((base + (q << log_trlo), col))
    return out


def run_group(I, inv, EG, log_r, log_c, log_trlo, nthreads, X, tw):
    for tid in range(nthreads):
        x = X[tid]
        for sg, (base, _) in enumerate(layout(EG, log_c, log_trlo, nthreads, tid)):
            hs = [1 << i for i in range(EG.bit_length() - 1)]
            for h in (hs if inv else hs[::-1]):
                lt = log_trlo + h.bit_length() - 1
                for q in range(EG):
                    if q & h:
                        continue
                    r = base + (q << log_trlo)
                    k = (1 << (log_r - lt - 1)) + (r >> (lt + 1))
                    assert 1 <= k < (1 << log_r)
                    u, v = x[sg * EG + q], x[sg * EG + q + h]
                    if inv:
                        x[sg * EG + q] = (u + v) % P[I]
                        x[sg * EG + q + h] = mont(I, (u - v) % P[I], tw[k])
                    else:
                        v = mont(I, v, tw[k])
                        x[sg * EG + q] = (u + v) % P[I]
                        x[sg * EG + q + h] = (u - v) % P[I]


def exchange(X, log_c, eg_from, tl_from, eg_to, tl_to, nthreads):
    lds = {}
    for tid in range(nthreads):
        for slot, (r, c) in enumerate(rows_cols(eg_from, log_c, tl_from, nthreads, tid)):
            a = (r << log_c) + c
            idx = a + (a >> 4)
            assert idx < KLDS and idx not in lds
            lds[idx] = X[tid][slot]
    for tid in range(nthreads):
        X[tid] = [lds[(((r << log_c) + c) + (((r << log_c) + c) >> 4))]
                  for (r, c) in rows_cols(eg_to, log_c, tl_to, nthreads, tid)]


def tile_transform(I, inv, X, log_r, log_c, tw, nthreads):
    gs = groups(log_r)
    order = range(len(gs) - 1, -1, -1) if inv else range(len(gs))
    prev = None
    for gi in order:
        st, tl = gs[gi]
        if prev is not None:
            exchange(X, log_c, 1 << gs[prev][0], gs[prev][1], 1 << st, tl, nthreads)
        run_group(I, inv, 1 << st, log_r, log_c, tl, nthreads, X, tw)
        prev = gi


def gio(plane, n, X, EG, log_c, log_trlo, tile_base, log_s, nthreads, store, limbs=None):
    for tid in range(nthreads):
        rc = rows_cols(EG, log_c, log_trlo, nthreads, tid)
        if not store:
            X[tid] = [0] * KE
        for slot, (r, c) in enumerate(rc):
            idx = tile_base + c + (r << log_s)
            assert 0 <= idx < n
            if store:
                plane[idx] = X[tid][slot]
            elif limbs is not None:
                X[tid][slot] = limbs[idx] if idx < len(limbs) else 0
            else:
                X[tid][slot] = plane[idx]


def plan(log_n):
    if log_n <= 12:
        return [], log_n
    q = log_n - 12
    rs = [q] if q <= 8 else [(q + 1) // 2, q // 2]
    outer, done = [], 0
    for r in rs:
        outer.append((r, 12 - r, log_n - done - r, 1 << done))
        done += r
    return outer, 12


def model_product(a, b, log_n):
    n = 1 << log_n
    outer, mid_log_r = plan(log_n)
    res = {0: [[0] * n for _ in range(3)], 1: [[0] * n for _ in range(3)]}
    ops = [(0, a), (1, b)]
    for I in range(3):
        Wf, Wi = wtab(I, False), wtab(I, True)
        for k, (log_r, log_c, log_s, nblk) in enumerate(outer):
            bt = beta_table(I, log_n, log_s, nblk, False)
            ntiles = 1 << (log_n - log_r - log_c)
            log_ncg = log_s - log_c
            gs = groups(log_r)
            for op, limbs in ops:
                for t in range(ntiles):
                    blk, cg = t >> log_ncg, t & ((1 << log_ncg) - 1)
                    tile_base = (blk << (log_r + log_s)) + (cg << log_c)
                    tw = build_tw(I, Wf, None if blk == 0 else bt[blk], log_r)
                    X = [None] * 256
                    gio(res[op][I], n, X, 1 << gs[0][0], log_c, gs[0][1], tile_base, log_s, 256,
                        False, limbs if k == 0 else None)
                    tile_transform(I, False, X, log_r, log_c, tw, 256)
                    gio(res[op][I], n, X, 1 << gs[-1][0], log_c, 0, tile_base, log_s, 256, True)
        # middle pass
        nblk = n >> mid_log_r
        bf = beta_table(I, log_n, 0, nblk, False)
        bi = beta_table(I, log_n, 0, nblk, True)
        nthreads = (1 << mid_log_r) // KE
        gs = groups(mid_log_r)
        for blk in range(nblk):
            tb = blk << mid_log_r
            twf = build_tw(I, Wf, None if blk == 0 else bf[blk], mid_log_r)
            twi = build_tw(I, Wi, None if blk == 0 else bi[blk], mid_log_r)
            XA, XB = [None] * nthreads, [None] * nthreads
            gio(res[0][I], n, XA, 1 << gs[0][0], 0, gs[0][1], tb, 0, nthreads, False,
                a if not outer else None)
            tile_transform(I, False, XA, mid_log_r, 0, twf, nthreads)
            gio(res[1][I], n, XB, 1 << gs[0][0], 0, gs[0][1], tb, 0, nthreads, False,
                b if not outer else None)
            tile_transform(I, False, XB, mid_log_r, 0, twf, nthreads)
            for t in range(nthreads):
                XA[t] = [mont(I, u, v) for u, v in zip(XA[t], XB[t])]
            tile_transform(I, True, XA, mid_log_r, 0, twi, nthreads)
            gio(res[0][I], n, XA, 1 << gs[0][0], 0, gs[0][1], tb, 0, nthreads, True)
        for k in range(len(outer) - 1, -1, -1):
            log_r, log_c, log_s, nblk = outer[k]
            bt = beta_table(I, log_n, log_s, nblk, True)
            ntiles = 1 << (log_n - log_r - log_c)
            log_ncg = log_s - log_c
            gs = groups(log_r)
            for t in range(ntiles):
                blk, cg = t >> log_ncg, t & ((1 << log_ncg) - 1)
                tile_base = (blk << (log_r + log_s)) + (cg << log_c)
                tw = build_tw(I, Wi, None if blk == 0 else bt[blk], log_r)
                X = [None] * 256
                gio(res[0][I], n, X, 1 << gs[-1][0], log_c, 0, tile_base, log_s, 256, False)
                tile_transform(I, True, X, log_r, log_c, tw, 256)
                gio(res[0][I], n, X, 1 << gs[0][0], log_c, gs[0][1], tile_base, log_s, 256, True)
    # Garner with the n^-1 R factor, pieces, and the map carry scan done sequentially
    y = res[0]
    out_len = len(a) + len(b)
    pieces = []
    for j in range(n):
        c = [mont(I, y[I][j], pow(n, -1, P[I]) * R32 * R32 % P[I]) for I in range(3)]
        v0 = c[0]
        v1 = (c[1] - v0) * pow(P[0], -1, P[1]) % P[1]
        t = (c[2] - v0) * pow(P[0], -1, P[2]) % P[2]
        v2 = (t - v1) * pow(P[1], -1, P[2]) % P[2]
        x = v0 + P[0] * (v1 + P[1] * v2)
        assert x < 1 << 84
        pieces.append((x & MASK, (x >> 28) & MASK, (x >> 56) & MASK))
    out, carry = [], 0
    for i in range(out_len):
        s = (pieces[i][0] if i < n else 0) + (pieces[i - 1][1] if 1 <= i <= n else 0) + \
            (pieces[i - 2][2] if 2 <= i <= n + 1 else 0)
        assert s < 3 << 28
        v = s + carry
        out.append(v & MASK)
        carry = v >> 28
        assert carry <= 3
    assert carry == 0
    return out


def tlog(la, lb):
    lg = 4
    while (1 << lg) < la + lb - 1:
        lg += 1
    return lg


@pytest.mark.parametrize("la,lb", [(1, 1), (3, 5), (8, 9), (100, 60), (1000, 1000), (2048, 2049)])
def test_single_block_schedule(la, lb):
    rng = random.Random(la * 7 + lb)
    a = [rng.randrange(1 << 28) for _ in range(la)]
    b = [rng.randrange(1 << 28) for _ in range(lb)]
    a[-1] |= 1
    b[-1] |= 1
    got = model_product(a, b, tlog(la, lb))
    exp = L.limbs_to_int(__import__("numpy").array(a, "uint32")) * \
        L.limbs_to_int(__import__("numpy").array(b, "uint32"))
    assert L.limbs_to_int(__import__("numpy").array(got, "uint32")) == exp


@pytest.mark.slow
@pytest.mark.parametrize("la,lb,log_n", [(3000, 2000, 13), (200, 30000, 15)])
def test_outer_pass_schedule(la, lb, log_n):
    """Transform lengths above one 4096-point tile: one outer pass plus the middle pass."""
    import numpy as np
    rng = random.Random(log_n)
    a = [rng.randrange(1 << 28) for _ in range(la)]
    b = [(1 << 28) - 1] * lb  # all-ones operand maximises coefficients
    a[-1] |= 1
    got = model_product(a, b, log_n)
    assert L.limbs_to_int(np.array(got, "uint32")) == \
        L.limbs_to_int(np.array(a, "uint32")) * L.limbs_to_int(np.array(b, "uint32"))


def test_plan_shapes():
    """Pass shapes for every supported length: tiles exactly cover n, C <= S, stages add up."""
    for log_n in range(4, 27):
        outer, mid = plan(log_n)
        assert sum(r for r, _, _, _ in outer) + mid == log_n
        for log_r, log_c, log_s, nblk in outer:
            assert log_r + log_c == 12 and log_c <= log_s
            assert nblk * (1 << (log_r + log_s)) == 1 << log_n
            assert 1 <= log_r <= 8 or (log_n - 12 > 8 and log_r <= 7)
