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

# The device primes and generators: the reference's (include/bigint/ntt_base.h:14-16).  P0 and P1
# are below 2^31 (room for lazy butterfly values); P2 = 2281701377 is above it.
P = [469762049, 2013265921, 2281701377]
G = [3, 31, 3]


def lazy_prime(I):
    return P[I] < (1 << 31)
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


KBETA = 16  # words per block in the beta power tables (kBetaPowStride)


def beta_table(I, log_n, log_s, nblk, inv, log_r):
    """Host upload_beta: per block b, beta_b = w_{n/S}^{bitrev(b)} as beta_b^(2^e), e < log_r,
    in Montgomery form, KBETA words per block."""
    if nblk <= 1:
        return None
    assert log_r <= KBETA
    lb = nblk.bit_length() - 1
    w = root(I, log_n - log_s, inv)
    return [[to_mont(I, pow(w, bitrev(b, lb) << e, P[I])) for e in range(log_r)] + [0] * (KBETA - log_r)
            for b in range(nblk)]


def build_tw(I, W, bpow, log_r):
    """Kernel build_twiddles: tw[k] = W[k] * beta^(2^(log_r - 1 - floor(log2 k)))."""
    R = 1 << log_r
    if bpow is None:
        return list(W[:R])
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
                    assert 0 <= u < P[I] and 0 <= v < P[I], "butterfly input out of range"
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


def word_residue(I, x):
    """Kernel word_residue: conditional subtractions bringing a 32-bit word below P (below
    4P for P0, whose forward butterflies are lazy)."""
    def csub(v, m):
        return v - m if v >= m else v
    if I == 0:
        x = csub(csub(x, 4 * P[0]), 4 * P[0])
        assert x < 4 * P[0]
    elif lazy_prime(I):
        x = csub(csub(x, 2 * P[I]), P[I])
        assert x < P[I]
    else:
        assert x < 2 * P[I]  # 32-bit word below 2 P2
        x = csub(x, P[I])
        assert x < P[I]
    return x % P[I]


def gio(plane, n, X, EG, log_c, log_trlo, tile_base, log_s, nthreads, store, limbs=None, prime=None):
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
                X[tid][slot] = word_residue(prime, limb_word(limbs, idx))
            else:
                X[tid][slot] = plane[idx]


M32 = 0xFFFFFFFF


def word_limb(g):
    """Kernel word_limb in 32-bit arithmetic: k = g + g / 7 (= floor(32 g / 28))."""
    assert 0 <= g <= M32
    k = g + g // 7
    assert k <= M32, "limb index overflows 32 bits"
    return k


def word_shift(g, k):
    """Kernel word_shift: 32 g - 28 k computed modulo 2^32."""
    return (g * 32 - k * 28) & M32


def limb_word(limbs, g):
    """Kernel limb_word: bits [32g, 32g+32) of the limb string, from limbs k and k+1."""
    k = word_limb(g)
    o = word_shift(g, k)
    assert o in (0, 4, 8, 12, 16, 20, 24)
    lo = limbs[k] if k < len(limbs) else 0
    hi = limbs[k + 1] if k + 1 < len(limbs) else 0
    return ((lo >> o) | (hi << (28 - o))) & 0xFFFFFFFF


def words_of(nlimbs):
    return (nlimbs * 7 + 7) // 8


KWT, KLT = 3584, 4096  # words / limbs per carry tile
GT, GW, GL = 512, 7, 8  # Garner threads per tile, words and limbs per thread


def map_apply(m, c):
    return (m >> (2 * c)) & 3


def map_compose(second, first):
    return sum(map_apply(second, map_apply(first, c)) << (2 * c) for c in range(4))


FIXUPS = []  # (tile, carry-in) of every tile the modelled fix-up changed


def garner_carry(coeff, out_len):
    """Kernels k_garner_carry and k_carry_fixup on exact coefficients coeff(j) (zero outside the
    product): 32-bit pieces, word sums with the two-word halo, per-thread carry maps over 7
    words, composition across threads, 7 words -> 8 limbs.  A tile's carry-in is 0 for tile 0
    and the local rule's value where it applies; otherwise the tile is written with carry-in 0
    and flagged, and the fix-up scans the tile maps and ripples the true carry-in into it.  The
    result is checked against the same arithmetic with every tile's true carry-in."""
    ntiles = (out_len + KLT - 1) // KLT
    out = [None] * out_len       # every tile with its true carry-in
    got = [None] * out_len       # the kernels' way
    tile_maps, flagged = [], []
    carry_in = 0
    for tile in range(ntiles):
        base = tile * KWT
        def piece(j, k):
            return (coeff(j) >> (32 * k)) & 0xFFFFFFFF if j >= 0 else 0
        sums = []
        for o in range(KWT):
            s = piece(base + o, 0) + piece(base + o - 1, 1) + piece(base + o - 2, 2)
            assert s < 3 << 32
            sums.append(s)
        # the kernel's sums: each thread's own coefficients j0 .. j0 + 6, plus the left
        # neighbour's w1(j0 - 1), w2(j0 - 1), w2(j0 - 2) (lane shuffle, LDS across waves; the
        # tile's halo coefficients base - 1, base - 2 for thread 0)
        for t in range(GT):
            j0 = base + GW * t
            own = []
            acc1 = acc2 = 0
            for q in range(GW):
                w0, w1, w2 = (piece(j0 + q, k) for k in range(3))
                own.append(acc1 + w0)
                acc1, acc2 = acc2 + w1, w2
            nb = j0 - 1  # the neighbour's last word (thread t - 1, or the halo for t = 0)
            own[0] += piece(nb, 1) + piece(nb - 1, 2)
            own[1] += piece(nb, 2)
            assert own == sums[GW * t:GW * t + GW], (tile, t)
        written = 0  # the carry-in the Garner kernel writes the tile with
        flag = False
        if tile:  # kernel's local resolution of the carry into the tile, checked against the scan
            sl = piece(base - 1, 0) + piece(base - 2, 1) + piece(base - 3, 2)
            if sl & 0xFFFFFFFF <= 0xFFFFFFFD:
                assert sl >> 32 == carry_in
                written = sl >> 32
            else:
                flag = True
        flagged.append(flag)
        maps = []
        for t in range(GT):
            c = [0, 1, 2, 3]
            for q in range(GW):
                s = sums[GW * t + q]
                c = [(s + ck) >> 32 for ck in c]
            assert max(c) <= 3 and c[0] <= c[1] <= c[2]  # carry-out non-decreasing in carry-in
            maps.append(c[0] | c[1] << 2 | c[2] << 4 | c[3] << 6)
        # the kernel's fast path: every run passes on a fixed carry (carry-outs for carry-ins 0
        # and 2 agree), so run t receives run t - 1's carry-out for carry-in 0
        co0 = [map_apply(m, 0) for m in maps]
        fixed_runs = all(map_apply(m, 0) == map_apply(m, 2) for m in maps)
        for cin, dst in ((carry_in, out), (written, got)):
            excl = 0b11100100
            for t in range(GT):
                carry = map_apply(excl, cin)
                assert carry <= 2
                if fixed_runs:
                    assert carry == (cin if t == 0 else co0[t - 1])
                words = []
                for q in range(GW):
                    v = sums[GW * t + q] + carry
                    words.append(v & 0xFFFFFFFF)
                    carry = v >> 32
                for q in range(GL):
                    bit = 28 * q
                    wi, bo = bit >> 5, bit & 31
                    v = words[wi] >> bo
                    if bo > 4:
                        v |= words[wi + 1] << (32 - bo)
                    i = tile * KLT + GL * t + q
                    if i < out_len:
                        dst[i] = v & MASK
                excl = map_compose(maps[t], excl)
        tile_maps.append(excl)
        carry_in = map_apply(excl, carry_in)
    assert carry_in == 0
    # k_carry_fixup: carry_in(0) = 0, carry_in(t + 1) = map_t(carry_in(t)); a flagged tile with a
    # nonzero carry-in gets it added at limb 0, rippling through the all-ones limbs of the tile
    c = 0
    for tile in range(ntiles):
        if flagged[tile] and c:
            FIXUPS.append((tile, c))
            lo, lim = tile * KLT, min(KLT, out_len - tile * KLT)
            v0 = got[lo] + c
            if v0 > MASK:
                first = next((i for i in range(1, lim) if got[lo + i] != MASK), lim)
                for i in range(1, first):
                    got[lo + i] = 0
                if first < lim:
                    got[lo + first] += 1
            got[lo] = v0 & MASK
        c = map_apply(tile_maps[tile], c)
    assert got == out
    return got


def coeff_pos(j, n):
    """Kernel coeff_pos: position of coefficient j in an inverse output computed with the forward
    roots (which leaves n x[(n - j) mod n] at position j)."""
    assert 0 <= j < n
    return (n - j) & (n - 1)


def load_pair_rev(plane, n, wx):
    """Kernel load_pair_rev: words wx and wx - 1 (wx even) of a plane, zero outside [0, n).  One
    unconditional load of the aligned position pair at n - wx for 2 <= wx <= n - 2, else at 0,
    then selects; the load index is bounds-checked whatever wx is."""
    assert wx % 2 == 0
    mid = 2 <= wx <= n - 2
    q = n - wx if mid else 0
    assert q % 2 == 0 and 0 <= q and q + 1 < n
    x, y = plane[q], plane[q + 1]
    return (x if mid or wx == 0 else 0), (y if mid or wx == n else 0)


def garner_stage(plane, n, base, chunks=None):
    """Kernel k_garner_carry's staging of the window [base - 4, base + 3584) of one plane: thread
    tid loads pairs i = k * 512 + tid (k < 4, i < 1792), threads 0, 1, 2 the pairs -2, -1, 1792;
    pair i holds words base + 2i and base + 2i - 1 and lands at staged indices 2i + 4 and 2i + 3.
    chunks: list of (plane, first word) summed per word (chunked products), else one plane at 0."""
    KST = KWT + 4
    s_res = [None] * KST
    srcs = chunks if chunks is not None else [(plane, 0)]

    def pair(i):
        x = y_ = 0
        for pl, c0 in srcs:
            a, b = load_pair_rev(pl, n, base + 2 * i - c0)
            x, y_ = x + a, y_ + b
        return x, y_
    for tid in range(GT):
        for k in range(4):
            i = k * GT + tid
            if i >= KWT // 2:
                for pl, _ in srcs:  # lanes past the tile's pairs load at wx = -2 (zeros)
                    assert load_pair_rev(pl, n, -2) == (0, 0)
                continue
            x, y_ = pair(i)
            t = 2 * i + 3
            assert s_res[t] is None and s_res[t + 1] is None
            s_res[t], s_res[t + 1] = y_, x
    x, _ = pair(-2)
    s_res[0] = x
    x, y_ = pair(-1)
    s_res[1], s_res[2] = y_, x
    _, y_ = pair(KWT // 2)
    assert s_res[KST - 1] is None
    s_res[KST - 1] = y_
    assert all(v is not None for v in s_res)
    return s_res


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
        Wf = wtab(I, False)
        for k, (log_r, log_c, log_s, nblk) in enumerate(outer):
            bt = beta_table(I, log_n, log_s, nblk, False, log_r)
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
                        False, limbs if k == 0 else None, I)
                    tile_transform(I, False, X, log_r, log_c, tw, 256)
                    gio(res[op][I], n, X, 1 << gs[-1][0], log_c, 0, tile_base, log_s, 256, True)
        # middle pass: one forward table per tile serves both directions
        nblk = n >> mid_log_r
        bf = beta_table(I, log_n, 0, nblk, False, mid_log_r)
        nthreads = (1 << mid_log_r) // KE
        gs = groups(mid_log_r)
        for blk in range(nblk):
            tb = blk << mid_log_r
            twf = build_tw(I, Wf, None if blk == 0 else bf[blk], mid_log_r)
            XA, XB = [None] * nthreads, [None] * nthreads
            gio(res[0][I], n, XA, 1 << gs[0][0], 0, gs[0][1], tb, 0, nthreads, False,
                a if not outer else None, I)
            tile_transform(I, False, XA, mid_log_r, 0, twf, nthreads)
            gio(res[1][I], n, XB, 1 << gs[0][0], 0, gs[0][1], tb, 0, nthreads, False,
                b if not outer else None, I)
            tile_transform(I, False, XB, mid_log_r, 0, twf, nthreads)
            for t in range(nthreads):
                XA[t] = [mont(I, u, v) for u, v in zip(XA[t], XB[t])]
            tile_transform(I, True, XA, mid_log_r, 0, twf, nthreads)
            gio(res[0][I], n, XA, 1 << gs[0][0], 0, gs[0][1], tb, 0, nthreads, True)
        for k in range(len(outer) - 1, -1, -1):
            log_r, log_c, log_s, nblk = outer[k]
            bt = beta_table(I, log_n, log_s, nblk, False, log_r)  # forward roots
            ntiles = 1 << (log_n - log_r - log_c)
            log_ncg = log_s - log_c
            gs = groups(log_r)
            for t in range(ntiles):
                blk, cg = t >> log_ncg, t & ((1 << log_ncg) - 1)
                tile_base = (blk << (log_r + log_s)) + (cg << log_c)
                tw = build_tw(I, Wf, None if blk == 0 else bt[blk], log_r)
                X = [None] * 256
                gio(res[0][I], n, X, 1 << gs[-1][0], log_c, 0, tile_base, log_s, 256, False)
                tile_transform(I, True, X, log_r, log_c, tw, 256)
                gio(res[0][I], n, X, 1 << gs[0][0], log_c, gs[0][1], tile_base, log_s, 256, True)
    # Garner with the n^-1 R factor, 32-bit pieces and the tile carry scan.  The inverse ran with
    # the forward roots, so coefficient j sits at position (n - j) mod n; each tile's window is
    # read through the kernel's pair staging.
    y = res[0]
    out_len = len(a) + len(b)
    for tile in range((out_len + KLT - 1) // KLT):
        for p in range(3):
            stage = garner_stage(y[p], n, tile * KWT)
            assert stage == [y[p][coeff_pos(w, n)] if 0 <= w < n else 0
                             for w in range(tile * KWT - 4, tile * KWT + KWT)]
    xs = []
    for j in range(n):
        pj = coeff_pos(j, n)
        x = kernel_crt(y[0][pj], y[1][pj], y[2][pj], n)
        assert x < 1 << 90
        xs.append(x)
    wa, wb = words_of(len(a)), words_of(len(b))
    wrap = limb_word(a, wa - 1) * limb_word(b, wb - 1) if wa + wb - 1 > n else 0

    def coeff(j):  # kernel garner(): the wrapped top coefficient leaves index 0 for index n
        if j >= n:
            return wrap if j == n else 0
        return xs[j] - wrap if j == 0 else xs[j]
    return garner_carry(coeff, out_len)


def tlog(la, lb):
    wa, wb = words_of(la), words_of(lb)
    need = max(wa + wb - 2, wa, wb)
    lg = 4
    while (1 << lg) < need:
        lg += 1
    return lg


@pytest.mark.parametrize("la,lb", [(1, 1), (3, 5), (8, 9), (100, 60), (1000, 1000), (2048, 2049),
                                   (2341, 2341), (19, 1), (18, 1)])
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


def _words(limbs):
    return [limb_word(limbs, g) for g in range(words_of(len(limbs)))]


@pytest.mark.parametrize("nlimbs", [1, 2, 7, 8, 9, 15, 16, 17, 100])
def test_limb_words_reassemble_the_integer(nlimbs):
    rng = random.Random(nlimbs)
    limbs = [rng.randrange(1 << 28) for _ in range(nlimbs)]
    limbs[-1] |= 1 << 27
    w = _words(limbs)
    assert sum(x << (32 * g) for g, x in enumerate(w)) == sum(x << (28 * i) for i, x in enumerate(limbs))
    assert w[-1] != 0 and all(0 <= x < 1 << 32 for x in w)
    # one word past the end reads zero
    assert limb_word(limbs, len(w)) == 0


@pytest.mark.parametrize("nlimbs", [700_000_000, 613_566_757, 0xFFFFFFFF - 8192])
def test_word_index_math_at_long_operands(nlimbs):
    """The first pass's word -> limb index math in the kernels' 32-bit arithmetic, replayed for
    operands far beyond 2^29 words (where 8 g overflows 32 bits): for words sampled over the whole
    operand and every word of its last 7-word groups, k = g + g / 7 and o = 32 g - 28 k (wrapping)
    equal the exact floor(32 g / 28) and its remainder, and k, k + 1 stay in 32 bits."""
    import numpy as np
    nw = words_of(nlimbs)
    rng = np.random.default_rng(nlimbs)
    g = np.concatenate([rng.integers(0, nw, 200_000, dtype=np.int64),
                        np.arange(max(0, nw - 4096), nw, dtype=np.int64),
                        np.arange(0, 4096, dtype=np.int64), [(1 << 29) - 1, 1 << 29]]).astype(np.int64)
    g = g[g < nw]
    gu = g.astype(np.uint32)
    k = gu + gu // np.uint32(7)                                   # word_limb, uint32
    o = gu * np.uint32(32) - k * np.uint32(28)                    # word_shift, wrapping uint32
    exact_k = (32 * g) // 28
    assert np.array_equal(k.astype(np.int64), exact_k)
    assert np.array_equal(o.astype(np.int64), 32 * g - 28 * exact_k)
    assert set(np.unique(o).tolist()) <= {0, 4, 8, 12, 16, 20, 24}
    assert int(k.max()) + 1 <= M32 and int(k.max()) < nlimbs
    # the formula it replaces, (8 g) / 7 in 32 bits, is wrong from g = 2^29 on
    old = (gu * np.uint32(8)) // np.uint32(7)
    assert (nw <= 1 << 29) or not np.array_equal(old.astype(np.int64), exact_k)
    # tile counts of the product stay exact in 32 bits at the largest supported product
    out_len = min(2 * nlimbs, 0xFFFFFFFF - 8192)
    assert out_len + 4095 <= M32 and (out_len + 4095) // 4096 * 3584 <= M32


@pytest.mark.parametrize("la,lb", [(1, 1), (8, 9), (4096, 3), (5000, 5000), (9000, 4)])
def test_garner_tile_carry_all_ones(la, lb):
    """Maximal coefficients: all-ones operands through the 32-bit-word tile carry."""
    a, b = (1 << (28 * la)) - 1, (1 << (28 * lb)) - 1
    wa = _words([MASK] * la)
    wb = _words([MASK] * lb)
    conv = [0] * (len(wa) + len(wb) - 1)
    # all-ones words except the top: convolution by prefix sums
    for i, x in enumerate(wa):
        for j, y in enumerate(wb):
            conv[i + j] += x * y
    out = garner_carry(lambda j: conv[j] if 0 <= j < len(conv) else 0, la + lb)
    got = sum(x << (28 * i) for i, x in enumerate(out))
    assert got == a * b


@pytest.mark.parametrize("la,lb,L,log_n", [
    (9000, 700, 1024, 11),      # power-of-two chunks: primary chunk j >> log_chunk
    (20000, 5000, 7168, 14),    # multiple of the 3584-word tile: one primary chunk per tile
    (20000, 9000, 7168, 14),    # shorter operand longer than the chunk: three chunks overlap
])
def test_chunked_garner_index_logic(la, lb, L, log_n):
    """The kernel's chunk loop (primary chunk, earlier chunks while off < n) summing exact chunk
    convolutions reproduces the full product."""
    rng = random.Random(L + la)
    A = [rng.randrange(1 << 28) for _ in range(la)]
    B = [rng.randrange(1 << 28) for _ in range(lb)]
    wa, wb = _words(A), _words(B)
    n = 1 << log_n
    assert L + len(wb) - 1 <= n
    nch = (len(wa) + L - 1) // L
    conv = []
    for c in range(nch):
        ch = wa[c * L:(c + 1) * L]
        v = [0] * n
        for i, x in enumerate(ch):
            if x:
                for j, y in enumerate(wb):
                    v[i + j] += x * y
        conv.append(v)
    pow2 = L & (L - 1) == 0
    log_chunk = L.bit_length() - 1

    def coeff(j):
        if j < 0:
            return 0
        c = (j >> log_chunk) if pow2 else (j // KWT * KWT) // L
        x = 0
        for d in range(c + 1):
            cc = c - d
            off = j - cc * L
            if off >= n:
                break
            if cc < nch:
                x += conv[cc][off]
        assert x < 1 << 89
        return x

    out = garner_carry(coeff, la + lb)
    got = sum(x << (28 * i) for i, x in enumerate(out))
    assert got == sum(x << (28 * i) for i, x in enumerate(A)) * sum(x << (28 * i) for i, x in enumerate(B))

    # the kernel's run form: for a thread's 16 coefficients from jl (even), chunks
    # (jl - n) / L < cc <= (jl + 15) / L, each adding at offsets 0 <= off < n
    total_words = len(wa) + len(wb)
    for jl in range(0, total_words + 64, 2):
        top = jl + 15
        c_hi = min(top // L, nch - 1)
        c_lo = (jl - n) // L + 1 if jl >= n else 0
        run = [0] * 16
        for cc in range(c_lo, c_hi + 1):
            off0 = jl - cc * L
            for i in range(0, 16, 2):
                off = off0 + i
                if 0 <= off < n:
                    run[i] += conv[cc][off]
                    run[i + 1] += conv[cc][off + 1]
        assert run == [coeff(jl + i) for i in range(16)], jl


def test_garner_carry_run_across_tiles():
    """A carry born in word 2 runs through three whole tiles (a * (2^32 + 1) with a_j + a_(j-1)
    = 2^32 - 1): the tile maps depend on the carry-in and must compose back to tile 0."""
    W = 3 * KWT + 100
    words = [(1 << 32) - 1, 5] + [((1 << 32) - 6) if j % 2 == 0 else 5 for j in range(2, W)]
    conv = [words[j] + (words[j - 1] if j else 0) for j in range(W)] + [words[-1]]
    a = sum(w << (32 * j) for j, w in enumerate(words))
    la = (32 * W + 27) // 28
    FIXUPS.clear()
    out = garner_carry(lambda j: conv[j] if 0 <= j < len(conv) else 0, la + 2)
    assert sum(x << (28 * i) for i, x in enumerate(out)) == a * ((1 << 32) + 1)
    # tiles 1-3 cannot resolve their carry-in locally; the fix-up adds it, rippling through them
    assert [t for t, _ in FIXUPS] == [1, 2, 3]


M32 = (1 << 32) - 1


def k_mont(I, a, b, lazy=False):
    """Kernel mont_mul (lazy: mont_mul_lazy) on 32-bit words, asserting the ranges it relies on:
    the 64-bit sum t + u*P must not overflow, and the shifted value must lie below 2P."""
    assert 0 <= a <= M32 and 0 <= b < P[I]
    t = a * b
    u = (t * (P[I] - 2)) & M32  # -P^-1 mod 2^32
    s = t + u * P[I]
    assert s < (1 << 64) and s % R32 == 0
    r = s >> 32
    assert r < 2 * P[I]
    return r if lazy or r < P[I] else r - P[I]


def k_redc(I, t):
    """Kernel mont_redc: Montgomery reduction of a 64-bit value, asserting its bounds (t below
    2^32 P and below 2^64 - 2^32 P, so t + u*P fits 64 bits and the shifted value is below 2P)."""
    assert 0 <= t < (1 << 32) * P[I] and t < (1 << 64) - (1 << 32) * P[I]
    u = ((t & M32) * (P[I] - 2)) & M32
    s = t + u * P[I]
    assert s < (1 << 64) and s % R32 == 0
    r = s >> 32
    assert r < 2 * P[I]
    r = r if r < P[I] else r - P[I]
    assert r == t * pow(R32, -1, P[I]) % P[I]
    return r


def k_add_mod(I, a, b):
    """Kernel add_mod in 32-bit words: for P < 2^31 the sum fits and an unsigned min of s and
    s - P selects; for P2 the carry out of 32 bits and the borrow of s - P decide."""
    assert 0 <= a < P[I] and 0 <= b < P[I]
    if lazy_prime(I):
        s = a + b
        assert s <= M32
        s2 = (s - P[I]) & M32
        return min(s, s2)
    s = (a + b) & M32
    carry = a + b > M32
    d = (s - P[I]) & M32
    borrow = s < P[I]
    r = d if (carry or not borrow) else s
    assert r == (a + b) % P[I]
    return r


def k_sub_mod(I, a, b):
    assert 0 <= a < P[I] and 0 <= b < P[I]
    if lazy_prime(I):
        d = (a - b) & M32
        return min(d, (d + P[I]) & M32)
    d = (a - b) & M32
    r = (d + P[I]) & M32 if a < b else d
    assert r == (a - b) % P[I]
    return r


def k_fold2p(x):
    assert 0 <= x < 4 * P[0]
    return x - 2 * P[0] if x >= 2 * P[0] else x


def lazy_group(I, inv, ST, x, rng):
    """Kernel group_stages on one register group of 2^ST values with its lazy reductions
    (P0: Harvey butterflies, forward values in [0, 4P), inverse in [0, 2P); P1: forward
    outputs that feed the next stage's Montgomery operand stay in [0, 2P), the inverse difference
    goes to Montgomery as u - v + P; P2 > 2^31: every value reduced, additions with a carry test).
    Every intermediate must fit 32 bits; the values are checked against the exact butterflies
    mod P."""
    p = P[I]
    rinv = pow(R32, -1, p)
    exact = [v % p for v in x]
    for hi in range(ST):
        lh = hi if inv else ST - 1 - hi
        h = 1 << lh
        for q in range(1 << ST):
            if q & h:
                continue
            w = rng.choice([p - 1, p - 2, 1, rng.randrange(p)])
            u, v = x[q], x[q + h]
            eu, ev = exact[q], exact[q + h]
            if inv:
                exact[q], exact[q + h] = (eu + ev) % p, (eu - ev) * w * rinv % p
            else:
                exact[q], exact[q + h] = (eu + ev * w * rinv) % p, (eu - ev * w * rinv) % p
            if I == 0:
                if not inv:
                    assert u < 4 * p and v < 4 * p
                    u2 = k_fold2p(u)
                    vw = k_mont(0, v, w, lazy=True)
                    x[q], x[q + h] = u2 + vw, u2 + 2 * p - vw
                    assert x[q] < 4 * p and 0 < x[q + h] < 4 * p
                else:
                    assert u < 2 * p and v < 2 * p
                    x[q] = k_fold2p(u + v)
                    x[q + h] = k_mont(0, u + 2 * p - v, w, lazy=True)
            elif not inv:
                vw = k_mont(I, v, w)
                if lazy_prime(I) and hi + 1 < ST and (q & (h >> 1)):
                    assert u < p
                    x[q], x[q + h] = u + vw, u + p - vw
                    assert x[q] < 2 * p <= M32 and x[q + h] < 2 * p
                else:
                    x[q], x[q + h] = k_add_mod(I, u, vw), k_sub_mod(I, u, vw)
            else:
                x[q] = k_add_mod(I, u, v)
                if lazy_prime(I):
                    x[q + h] = k_mont(I, u + p - v, w)
                else:
                    x[q + h] = k_mont(I, k_sub_mod(I, u, v), w)
            assert x[q] % p == exact[q] and x[q + h] % p == exact[q + h]
    return x


@pytest.mark.parametrize("I", [0, 1, 2])
@pytest.mark.parametrize("inv", [False, True])
@pytest.mark.parametrize("ST", [1, 2, 3, 4])
def test_lazy_butterfly_ranges(I, inv, ST):
    rng = random.Random(1000 * I + 10 * ST + inv)
    p = P[I]
    top = (4 * p if not inv else 2 * p) if I == 0 else p  # input range of a group
    for _ in range(150):
        x = [rng.choice([top - 1, top - 2, 0, rng.randrange(top)]) for _ in range(1 << ST)]
        out = lazy_group(I, inv, ST, x, rng)
        assert all(v < top for v in out)


@pytest.mark.parametrize("I", [0, 1, 2])
def test_word_residue_and_pointwise_ranges(I):
    """Input words reduce into a forward group's input range, and the pointwise product of two
    forward outputs lands in an inverse group's input range."""
    p = P[I]
    for x in [0, 1, p - 1, p, 2 * p - 1, 2 * p, 4 * p - 1 if I == 0 else M32, M32]:
        if x > M32:  # inputs are 32-bit words
            continue
        r = word_residue(I, x)
        assert r == x % p
    fwd_top = 4 * p if I == 0 else p
    for a, b in [(fwd_top - 1, fwd_top - 1), (0, fwd_top - 1), (fwd_top // 3, fwd_top - 7)]:
        if I == 0:
            fa, fb = k_fold2p(a), k_fold2p(b)
            t = fa * fb  # mont_mul_lazy(fold2p(a), fold2p(b)): both below 2P
            u = (t * (p - 2)) & M32
            s = t + u * p
            assert s < (1 << 64)
            r = s >> 32
            assert r < 2 * p and r % p == a * b * pow(R32, -1, p) % p
        else:
            r = k_mont(I, a, b)
            assert r < p


def kernel_crt(r0, r1, r2, n):
    """Kernel crt_coeff: the residues carry the inverse transform's factor n * R^-1, removed by
    the Garner scale constants n^-1 * R^2; every modular step is checked for its input ranges."""
    ks = [pow(n, -1, P[I]) * R32 * R32 % P[I] for I in range(3)]
    inv01 = to_mont(1, pow(P[0], -1, P[1]))
    inv02 = to_mont(2, pow(P[0], -1, P[2]))
    inv12 = to_mont(2, pow(P[1], -1, P[2]))
    assert r0 < 2 * P[0] and r1 < P[1] and r2 < P[2]
    # crt_consts: the scale folded into the first Garner products, their factors negated
    a1, a2 = k_mont(1, ks[1], inv01), k_mont(2, ks[2], inv02)
    m01, m02 = P[1] - inv01, P[2] - inv02
    v0 = k_mont(0, r0, ks[0])
    # v1 = (c1 - v0) inv01 and t = (c2 - v0) inv02 (c_i = r_i n^-1) as one reduction each
    v1 = k_redc(1, r1 * a1 + v0 * m01)
    t = k_redc(2, r2 * a2 + v0 * m02)
    assert v1 == k_mont(1, k_sub_mod(1, k_mont(1, r1, ks[1]), v0), inv01)
    assert t == k_mont(2, k_sub_mod(2, k_mont(2, r2, ks[2]), v0), inv02)
    assert v1 < P[1] < P[2]  # already a residue mod P2
    v2 = k_mont(2, k_sub_mod(2, t, v1), inv12)
    x = v0 + P[0] * (v1 + P[1] * v2)
    assert x < P[0] * P[1] * P[2]
    return x


def test_crt_recovers_every_coefficient():
    """Coefficients up to the largest possible (2^26 terms below 2^64: under 2^90) come back from
    their three residues, including P0 residues given in [P0, 2 P0) as the lazy inverse leaves
    them."""
    rng = random.Random(7)
    assert P[0] * P[1] * P[2] > (1 << 26) * ((1 << 32) - 1) ** 2
    for I in (1, 2):  # mont_redc's bound at the largest two-product sums of the CRT
        t = (P[I] - 1) ** 2 + (P[0] - 1) * (P[I] - 1)
        assert t < (1 << 32) * P[I] and t < (1 << 64) - (1 << 32) * P[I]
    xs = [0, 1, P[0], P[0] * P[1], P[2] * P[1] - 1, (1 << 26) * ((1 << 32) - 1) ** 2]
    xs += [rng.randrange(1 << 90) for _ in range(3000)] + [rng.randrange(1 << 64) for _ in range(500)]
    for n in (2, 1 << 11, 1 << 24):
        rinv = [pow(R32, -1, P[I]) for I in range(3)]
        for x in xs:
            r = [x * n * rinv[I] % P[I] for I in range(3)]
            assert kernel_crt(*r, n) == x
            if r[0] < P[0]:
                assert kernel_crt(r[0] + P[0], r[1], r[2], n) == x


# ---- batched device additions (bigint_amd/csrc/add_kernels.hip) -------------------------------

ATILE, ATHREADS, ALIMBS = 4096, 256, 16
AIDENT = 0b10


def amap_apply(m, c):
    return (m >> c) & 1


def amap_compose(second, first):
    return amap_apply(second, amap_apply(first, 0)) | (amap_apply(second, amap_apply(first, 1)) << 1)


def astep(a, b, c, sub):
    if not sub:
        s = a + b + c
        return s >> 28, s & MASK
    t = b + c
    return (1 if a < t else 0), (a - t) & MASK


def athread_map(a, b, sub):
    c0, c1 = 0, 1
    for x, y in zip(a, b):
        c0, _ = astep(x, y, c0, sub)
        c1, _ = astep(x, y, c1, sub)
    return c0 | (c1 << 1)


def model_add(x, y, sub):
    """The three kernels of one job: tile maps, per-job scan, apply (with the thread-level
    exclusive scan inside a tile).  Returns (limbs, swapped, length)."""
    width = max(len(x), len(y)) + (0 if sub else 1)
    ntiles = -(-width // ATILE)
    lim = lambda v, i: v[i] if i < len(v) else 0

    def tile_limbs(t, first, second):
        base = t * ATILE
        return ([lim(first, base + i) for i in range(ATILE)], [lim(second, base + i) for i in range(ATILE)])

    def tile_total(a, b, s):
        tot = AIDENT
        for th in range(ATHREADS):
            tot = amap_compose(athread_map(a[th * ALIMBS:(th + 1) * ALIMBS], b[th * ALIMBS:(th + 1) * ALIMBS], s), tot)
        return tot

    maps = []
    for t in range(ntiles):
        a, b = tile_limbs(t, x, y)
        if not sub:
            maps.append(tile_total(a, b, False))
        else:
            maps.append(tile_total(a, b, True) | (tile_total(b, a, True) << 2))
    run0 = run1 = AIDENT
    carry_in = []
    for t in range(ntiles):
        carry_in.append(amap_apply(run0, 0) | (amap_apply(run1, 0) << 1))
        run0 = amap_compose(maps[t] & 3, run0)
        if sub:
            run1 = amap_compose((maps[t] >> 2) & 3, run1)
    swap = sub and amap_apply(run0, 0) == 1
    out = [None] * width
    top = 0
    for t in range(ntiles):
        a, b = tile_limbs(t, y, x) if swap else tile_limbs(t, x, y)
        c = (carry_in[t] >> (1 if swap else 0)) & 1
        excl = AIDENT
        for th in range(ATHREADS):
            ta, tb = a[th * ALIMBS:(th + 1) * ALIMBS], b[th * ALIMBS:(th + 1) * ALIMBS]
            cc = amap_apply(excl, c)
            for q in range(ALIMBS):
                cc, o = astep(ta[q], tb[q], cc, sub)
                i = t * ATILE + th * ALIMBS + q
                if i < width:
                    out[i] = o
                    if o:
                        top = max(top, i + 1)
            excl = amap_compose(athread_map(ta, tb, sub), excl)
    assert None not in out
    return out, swap, top


@pytest.mark.parametrize("lx,ly", [(1, 1), (3, 7), (4096, 4096), (4097, 5), (2, 8200)])
def test_device_add_model(lx, ly):
    rng = random.Random(lx * 10007 + ly)
    val = lambda v: sum(d << (28 * i) for i, d in enumerate(v))
    cases = [([rng.randrange(1 << 28) for _ in range(lx)], [rng.randrange(1 << 28) for _ in range(ly)])]
    cases.append(([MASK] * lx, [1] + [0] * (ly - 1)))              # carries through every limb
    cases.append(([0] * (lx - 1) + [1], [1] * min(ly, lx)))       # borrows through every limb
    same = [rng.randrange(1 << 28) for _ in range(lx)]
    cases.append((same, list(same)))
    for x, y in cases:
        for sub in (False, True):
            out, swap, n = model_add(x, y, sub)
            want = val(x) + val(y) if not sub else abs(val(x) - val(y))
            assert val(out) == want
            assert swap == (sub and val(y) > val(x))
            assert n == (want.bit_length() + 27) // 28  # 0 for zero


def _source_words(nlimbs, off, cnt):
    w = words_of(nlimbs)
    return min(max(w - off, 0), cnt)


@pytest.mark.parametrize("nlimbs,off,cnt,log_n,log_r", [
    (9586981 // 64, 0, M32, 18, 7),          # a 2^22-bit operand on its 2^18-point transform
    (4682, 0, M32, 13, 6),                    # 4097 words: one word into the upper half
    (2341, 0, M32, 12, 4),                    # 2049 words on 2^12 points (the wrapped case)
    (1000, 0, M32, 12, 8), (30, 0, M32, 6, 2), (5, 0, M32, 4, 1),
    (60000, 3 * 4096, 4096, 13, 5),           # a chunk of a longer operand
    (60000, 13 * 4096, 4096, 13, 7),          # the last, partial chunk
])
def test_first_pass_half_skip_reads_only_zero_rows(nlimbs, off, cnt, log_n, log_r):
    """The first forward pass skips the upper half rows of a tile (and replaces the first stage by
    a copy) when tile_base + (R/2) S >= the source's word count.  Replayed over every tile: each
    skipped element really is a zero input word, and the copy equals the butterfly it replaces."""
    rng = random.Random(nlimbs + off)
    limbs = [rng.randrange(1 << 28) for _ in range(nlimbs)]
    limbs[-1] |= 1 << 27
    log_s = log_n - log_r
    log_c = 12 - log_r
    if log_c > log_s:
        pytest.skip("tile wider than the row stride")
    sw = _source_words(nlimbs, off, cnt)
    skipped = 0
    for t in range(1 << (log_n - 12)):
        tile_base = t << log_c          # first pass: one block, column groups
        half = tile_base + (1 << (log_r - 1 + log_s)) >= sw
        if not half:
            continue
        skipped += 1
        for r in range(1 << (log_r - 1), 1 << log_r):
            for c in range(1 << log_c):
                idx = tile_base + c + (r << log_s)
                assert idx < (1 << log_n)
                word = limb_word(limbs, off + idx) if idx < cnt else 0
                assert word == 0, (t, r, c)
    assert skipped > 0 or sw > (1 << (log_n - 1))
    # the copy is the butterfly with v = 0 for every prime and twiddle
    for I in range(3):
        for u in [0, 1, P[I] - 1]:
            for w in [to_mont(I, 1), to_mont(I, 12345)]:
                assert (u + k_mont(I, 0, w)) % P[I] == u and (u - k_mont(I, 0, w)) % P[I] == u


@pytest.mark.parametrize("n", [2, 4, 8, 64, 4096])
def test_pair_load_matches_word_definition(n):
    """The branch-free pair load returns words wx and wx - 1 of the coefficient sequence (word w at
    position (n - w) mod n, zero outside [0, n)) for every even wx around and inside the plane,
    with its one load inside the plane for each of them."""
    plane = [1000 + q for q in range(n)]

    def word(w):
        return plane[coeff_pos(w, n)] if 0 <= w < n else 0
    for wx in range(-8, n + 10, 2):
        assert load_pair_rev(plane, n, wx) == (word(wx), word(wx - 1)), wx


@pytest.mark.parametrize("n", [8192, 16384, 1 << 16])
def test_interior_garner_loads_stay_in_the_plane(n):
    """Interior Garner tiles load unconditionally: lanes past the tile's 1792 pairs reload pair 0 and
    lanes without a halo pair load the tile's last pair (ie = 1792).  Every such position, for every
    lane of every interior tile, is an aligned pair inside the plane."""
    npairs = KWT // 2
    for base in range(0, n, KWT):
        if not (base >= 6 and base + KWT + 2 <= n):
            continue
        top = n - base
        for tid in range(GT):
            for k in range(4):
                i = k * GT + tid
                q = top - 2 * (i if i < npairs else 0)
                assert q % 2 == 0 and 0 <= q and q + 1 < n
            ie = -2 if tid == 0 else (-1 if tid == 1 else npairs)
            q = top - 2 * ie
            assert q % 2 == 0 and 0 <= q and q + 1 < n


@pytest.mark.parametrize("n,L,nch", [(16, 8, 5), (64, 32, 3), (4096, 2048, 4), (8192, 3584 * 2, 3),
                                     (8192, 4096, 5)])
def test_garner_staging_chunked_reversed(n, L, nch):
    """Chunked Garner staging: chunk cc's transform holds its word wl at position (n - wl) mod n;
    the kernel sums, per window word w, the chunks with 0 <= w - cc L < n, loading aligned
    position pairs.  Replayed for every tile of the product against the direct sum."""
    rng = random.Random(n * 7 + L + nch)
    planes = [[rng.randrange(1 << 20) for _ in range(n)] for _ in range(nch)]
    total = (nch - 1) * L + n
    for base in range(0, total + KWT, KWT):
        lo, hi = max(base - 4, 0), base + KWT - 1
        c_hi = min(hi // L, nch - 1)
        c_lo = (lo - n) // L + 1 if lo >= n else 0
        chunks = [(planes[cc], cc * L) for cc in range(c_lo, c_hi + 1)]
        got = garner_stage(None, n, base, chunks)
        exp = []
        for w in range(base - 4, base + KWT):
            v = 0
            for cc in range(nch):
                wl = w - cc * L
                if 0 <= wl < n:
                    v += planes[cc][coeff_pos(wl, n)]
            exp.append(v)
        assert got == exp, base


def tile_twiddles(I, W, bpow, log_r, nt):
    """TileTwiddles<I, LR, NT> (ntt_kernels.hip): thread tid builds entries k = tid + NT*j for
    j < J = max(1, 2^LR / NT); the beta power of j >= 1 is bpow[LR - 1 - log2 NT - floor(log2 j)],
    one scalar per j for the whole block; j = 0 takes bpow[LR - 1 - floor(log2 (tid | 1))]; with
    no per-block root every power is the Montgomery form of 1."""
    R = 1 << log_r
    J = R // nt if R >= nt else 1
    lnt, lj = nt.bit_length() - 1, J.bit_length() - 1
    one = R32 % P[I]
    tw = [None] * R
    for tid in range(nt):
        if R < nt and tid >= R:
            continue
        w = [W[tid + nt * j] for j in range(J)]
        if bpow is None:
            b0, bj = one, [one] * lj
        else:
            e0 = log_r - 1 - ((tid | 1).bit_length() - 1)
            assert 0 <= e0 < log_r
            b0 = bpow[e0]
            assert all(0 <= log_r - 1 - lnt - e < log_r for e in range(lj))
            bj = [bpow[log_r - 1 - lnt - e] for e in range(lj)]
        for j in range(J):
            assert tw[tid + nt * j] is None
            tw[tid + nt * j] = mont(I, w[j], b0 if j == 0 else bj[j.bit_length() - 1])
    return tw


@pytest.mark.parametrize("I", [0, 1, 2])
def test_tile_twiddles_match_the_entrywise_build(I):
    """The two-step per-thread table build equals the entry-by-entry definition
    tw[k] = W[k] * beta^(2^(LR - 1 - floor(log2 k))) for every entry the butterflies read (k >= 1),
    for the strided passes (LR 1..8, 256 threads) and the block passes (LR 4..12, 2^(LR-4) threads),
    with and without a per-block root."""
    rng = random.Random(40 + I)
    W = wtab(I, False)
    beta = rng.randrange(2, P[I])
    for log_r, nt in [(lr, 256) for lr in range(1, 9)] + [(lr, 1 << (lr - 4)) for lr in range(4, 13)]:
        bpow = [to_mont(I, pow(beta, 1 << e, P[I])) for e in range(log_r)]
        for bp in (bpow, None):
            got = tile_twiddles(I, W, bp, log_r, nt)
            want = build_tw(I, W, bp, log_r)
            assert got[1:] == want[1:], (log_r, nt, bp is None)
