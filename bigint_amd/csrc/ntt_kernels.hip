// Three-prime NTT multiply kernels for gfx950 (CDNA4), written for this engine.
//
// Algorithm (same mathematics as the reference's ntt::mul, src/mul/ntt_multithread.cpp:444-530,
// with a GPU-shaped schedule):
//   * residues live in planar arrays, one plane of n u32 per prime (the reference interleaves
//     three residues per point, ntt_multithread.cpp:447-483);
//   * forward transform = Cooley-Tukey stages, natural order in, bit-reversed order out, so no
//     bit-reversal permutation is ever done (the reference runs multi_bit_swap, ntt_base.h:268-287);
//   * the pointwise product happens in bit-reversed order and the inverse (Gentleman-Sande
//     stages) maps it back to natural order scaled by n;
//   * stages are grouped into passes.  A pass works on tiles of R rows x C columns at row stride
//     S in one workgroup: 16 elements per thread in registers, four radix-2 stages per register
//     group, groups exchanged through LDS;
//   * the last pass of each forward transform (contiguous 4096-point blocks) is fused with the
//     pointwise product and the first inverse pass;
//   * twiddles for a tile are generated into LDS from a 4096-entry root table and one per-block
//     root (tw[k] = W[k] * beta^tr), so no n-sized twiddle table is read;
//   * Garner CRT is fused with carry propagation: a scan of carry maps {0..3} -> {0..3} inside
//     each tile, the carry between tiles settled from the previous tile's last word, and a fix-up
//     pass that scans the tile maps for the rare tile that cannot settle it
//     (ntt_multithread.cpp:384-442 does the same CRT and a serial three-way overlapped carry).
//
// Twiddle convention: W[m + i] = w_{2m}^{bitrev_{log2 m}(i)} (Montgomery form), used by the
// butterfly of block i at the stage with m blocks, in the forward and in the inverse stages (the
// inverse then computes the transform with root w, which leaves coefficient j at position
// (n - j) mod n: see coeff_pos).  Data stay in plain residue form: a Montgomery
// product with a Montgomery-form twiddle is an ordinary modular product.  The pointwise product
// is a Montgomery product, so the inverse output carries n * R^-1, removed in Garner.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <array>
#include <utility>

#include "ntt_device.h"
#include "ntt_plan.h"

namespace bi {

constexpr int kE = 16;                        // elements per thread in every pass
constexpr int kTile = 4096;                   // elements per tile per prime
constexpr int kThreads = kTile / kE;          // 256
constexpr int kLdsWords = kTile + kTile / 16; // data tile with one pad word per 16

__device__ __forceinline__ uint32_t lds_idx(uint32_t a) { return a + (a >> 4); }

template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    [&]<int... Is>(std::integer_sequence<int, Is...>) {
        (f(std::integral_constant<int, Is>{}), ...);
    }(std::make_integer_sequence<int, N>{});
}

// Register-group schedule of a tile with 2^LR rows: ceil(LR/4) groups from the top rows down,
// four radix-2 stages each except the last, which takes the remainder.  Group gi covers the
// stages whose row half-span tr satisfies 2^trlo(gi) <= tr < 2^(trlo(gi) + stages(gi)).
template <int LR>
struct Sched {
    static constexpr int NG = (LR + 3) / 4;
    static constexpr int stages(int gi) { return (LR - 4 * gi) >= 4 ? 4 : (LR - 4 * gi); }
    static constexpr int trlo(int gi) { return LR - 4 * gi - stages(gi); }
};

// Tile of 2^LR rows x 2^LC columns, 16 elements per thread.  In a group with EG = 2^ST elements
// per register group at row stride 2^TL, sub-group sg of thread tid holds rows
// base + q * 2^TL (q < EG) of column col, where s = tid + T*sg is split as
// (rh | rl | col) = (s >> (LC + TL), (s >> LC) & (2^TL - 1), s & (2^LC - 1)) and
// base = rh * 2^(TL + ST) + rl.
template <int LR, int LC, int ST, int TL>
struct Layout {
    static constexpr int EG = 1 << ST;
    static constexpr int NSG = kE / EG;
    static constexpr uint32_t T = 1u << (LR + LC - 4);
    __device__ __forceinline__ static uint32_t s_of(int sg) { return threadIdx.x + T * (uint32_t)sg; }
    __device__ __forceinline__ static uint32_t col(int sg) { return s_of(sg) & ((1u << LC) - 1u); }
    __device__ __forceinline__ static uint32_t rh(int sg) { return s_of(sg) >> (LC + TL); }
    __device__ __forceinline__ static uint32_t row(int sg, int q) {
        const uint32_t s = s_of(sg);
        const uint32_t rl = (s >> LC) & ((1u << TL) - 1u);
        return ((s >> (LC + TL)) << (TL + ST)) + rl + ((uint32_t)q << TL);
    }
};

// Twiddle table of one tile: tw[k], k in [1, 2^LR), for the stage with half-span
// tr = 2^LR / (2 * 2^floor(log2 k)) rows:  tw[k] = W[k] * beta^tr.  bpow holds the block's
// beta^(2^e), e < LR (precomputed per block, kBetaPowStride words); null encodes beta = 1, for
// which the powers are the Montgomery form of 1 (a product with it returns the root unchanged, so
// every block takes the same branch-free build).
//
// Built in two steps by NT threads, entry k = tid + NT*j for j < 2^LR / NT: load() issues every
// load of the build at once, so a caller can put them ahead of its tile's data loads, and store()
// forms the products and writes the table, waiting on memory once (a loop that loads and
// multiplies entry by entry waits for each entry's loads, and for every load in flight before
// them).  For j >= 1, floor(log2 k) = log2 NT + floor(log2 j) is the same for the whole block, so
// those beta powers come from scalar loads; only j = 0 picks its power per lane.
template <int I, int LR, int NT>
struct TileTwiddles {
    static_assert(LR <= (int)kBetaPowStride, "beta power table too short");
    static_assert((NT & (NT - 1)) == 0, "thread count must be a power of two");
    static constexpr int R = 1 << LR;
    static constexpr int J = R >= NT ? R / NT : 1;  // entries per thread
    static constexpr int LNT = __builtin_ctz(NT);
    static constexpr int LJ = __builtin_ctz(J);     // distinct uniform powers (j >= 1)
    // 2^32 mod P: the Montgomery form of 1, the beta power of a block without a root
    static constexpr uint32_t kOne = (uint32_t)((1ull << 32) % kP[I]);
    uint32_t w[J];
    uint32_t b0;
    uint32_t bj[LJ > 0 ? LJ : 1];

    __device__ __forceinline__ void load(const uint32_t* __restrict__ wtab,
                                         const uint32_t* __restrict__ bpow) {
        const uint32_t tid = threadIdx.x;
#pragma unroll
        for (int j = 0; j < J; ++j) w[j] = (R >= NT || tid < (uint32_t)R) ? gmem(wtab)[tid + NT * j] : 0u;
        b0 = kOne;
#pragma unroll
        for (int e = 0; e < LJ; ++e) bj[e] = kOne;
        if (bpow) {
            const uint32_t log_m = 31u - __builtin_clz(tid | 1u);  // k = 0 is never read
            if (R >= NT || tid < (uint32_t)R) b0 = gmem(bpow)[LR - 1 - (int)log_m];
#pragma unroll
            for (int e = 0; e < LJ; ++e) bj[e] = gmem(bpow)[LR - 1 - LNT - e];
        }
    }

    __device__ __forceinline__ void store(uint32_t* __restrict__ tw) const {
        const uint32_t tid = threadIdx.x;
        if (R < NT && tid >= (uint32_t)R) return;
        tw[tid] = mont_mul<I>(w[0], b0);
#pragma unroll
        for (int j = 1; j < J; ++j) tw[tid + NT * j] = mont_mul<I>(w[j], bj[31 - __builtin_clz(j)]);
    }
};

// The block's beta powers in a pass's table ([prime][block][kBetaPowStride]), or null for
// block 0 (beta = 1) and for passes without per-block roots.
template <int I>
__device__ __forceinline__ const uint32_t* block_pows(const uint32_t* table, uint32_t stride,
                                                      uint32_t blk) {
    return (table && blk != 0) ? table + ((size_t)I * stride + blk) * kBetaPowStride : nullptr;
}

// Forward (Cooley-Tukey) or inverse (Gentleman-Sande) stages of one register group.
// Butterfly with lower row r at the stage with half-span 2^lt uses tw[2^(LR-lt-1) + (r >> (lt+1))].
// SKIP0 with skip0 set (forward only): the upper operand of every butterfly of the group's first
// stage is zero, so that stage leaves (u, u): a copy instead of butterflies.
template <int I, bool INV, int LR, int LC, int ST, int TL, bool SKIP0 = false>
__device__ __forceinline__ void group_stages(uint32_t (&x)[kE], const uint32_t* __restrict__ tw,
                                             bool skip0 = false) {
    using Ly = Layout<LR, LC, ST, TL>;
    constexpr int EG = Ly::EG;
    static_for<ST>([&](auto hi) {
        constexpr int lh = INV ? (int)hi : (ST - 1 - (int)hi);  // log2 of the span in q units
        constexpr int h = 1 << lh;
        constexpr int lt = TL + lh;
        if constexpr (SKIP0 && !INV && (int)hi == 0) {
            if (skip0) {
                // u + 0 * w and u - 0 * w: both outputs are u, which lies in the next stage's
                // input range for every prime (below 4P for P0, below P for P1 and P2)
#pragma unroll
                for (int sg = 0; sg < Ly::NSG; ++sg)
#pragma unroll
                    for (int q = 0; q < EG; ++q)
                        if (!(q & h)) x[sg * EG + q + h] = x[sg * EG + q];
                return;
            }
        }
#pragma unroll
        for (int sg = 0; sg < Ly::NSG; ++sg) {
            const uint32_t rhv = Ly::rh(sg);
#pragma unroll
            for (int q = 0; q < EG; ++q) {
                if (q & h) continue;
                const uint32_t k = (1u << (LR - lt - 1)) + (rhv << (TL + ST - lt - 1)) +
                                   ((uint32_t)q >> (lh + 1));
                const uint32_t w = tw[k];
                const uint32_t u = x[sg * EG + q];
                const uint32_t v = x[sg * EG + q + h];
                if constexpr (I == 0) {
                    // P0 < 2^29: Harvey-style lazy butterflies.  Forward values live in [0, 4P),
                    // inverse values in [0, 2P); both fold to [0, P) only where a consumer needs it.
                    constexpr uint32_t p2 = 2u * kP[0];
                    if (!INV) {
                        const uint32_t u2 = fold2p<0>(u);
                        const uint32_t vw = mont_mul_lazy<0>(v, w);
                        x[sg * EG + q] = u2 + vw;
                        x[sg * EG + q + h] = u2 + p2 - vw;
                    } else {
                        x[sg * EG + q] = fold2p<0>(u + v);
                        x[sg * EG + q + h] = mont_mul_lazy<0>(u + p2 - v, w);
                    }
                } else if (!INV) {
                    const uint32_t vw = mont_mul<I>(v, w);
                    // P1 < 2^31: outputs that are the multiplied operand ("v") of the next stage
                    // in this register group stay in [0, 2P) (Montgomery accepts any 32-bit
                    // input with P < 2^31); the rest, and every output of a group's last stage,
                    // are reduced.  P2 > 2^31: every output reduced.
                    constexpr bool kLazy = lazy_prime(I) && (int)hi + 1 < ST;
                    if (kLazy && (q & (h >> 1)) != 0) {
                        x[sg * EG + q] = u + vw;
                        x[sg * EG + q + h] = u + kP[I] - vw;
                    } else {
                        x[sg * EG + q] = add_mod<I>(u, vw);
                        x[sg * EG + q + h] = sub_mod<I>(u, vw);
                    }
                } else {
                    x[sg * EG + q] = add_mod<I>(u, v);
                    if constexpr (lazy_prime(I)) {
                        // u - v + P < 2P < 2^32 goes to Montgomery unreduced
                        x[sg * EG + q + h] = mont_mul<I>(u + kP[I] - v, w);
                    } else {
                        x[sg * EG + q + h] = mont_mul<I>(sub_mod<I>(u, v), w);
                    }
                }
            }
        }
    });
}

template <int LR, int LC, int ST, int TL, bool STORE>
__device__ __forceinline__ void lds_io(uint32_t* lds, uint32_t (&x)[kE]) {
    using Ly = Layout<LR, LC, ST, TL>;
#pragma unroll
    for (int sg = 0; sg < Ly::NSG; ++sg) {
        const uint32_t c = Ly::col(sg);
#pragma unroll
        for (int q = 0; q < Ly::EG; ++q) {
            const uint32_t a = lds_idx((Ly::row(sg, q) << LC) + c);
            if (STORE) lds[a] = x[sg * Ly::EG + q];
            else x[sg * Ly::EG + q] = lds[a];
        }
    }
}

// Limb index k = floor(32g / 28) = g + floor(g / 7) of transform word g, in 32-bit arithmetic
// without overflow: an operand has fewer than 2^32 limbs, so g < 7 * 2^29 and k < 2^32 (8g itself
// would overflow from g = 2^29 on).
__device__ __forceinline__ uint32_t word_limb(uint32_t g) { return g + g / 7u; }

// Bit offset o = 32g - 28k of word g inside limb k, in {0, 4, ..., 24}: the exact value is small,
// so the wrapping 32-bit difference is exact.
__device__ __forceinline__ uint32_t word_shift(uint32_t g, uint32_t k) { return g * 32u - k * 28u; }

// Transform input word g of an operand: bits [32g, 32g + 32) of its limb string.  The bits lie
// in limbs k = floor(32g / 28) and k + 1, from bit o = 32g - 28k (o in {0, 4, ..., 24}) of limb k;
// limbs at and past nlimbs read as zero.
__device__ __forceinline__ uint32_t limb_word(const uint32_t* __restrict__ limbs, uint32_t nlimbs,
                                              uint32_t g) {
    const uint32_t k = word_limb(g);
    const uint32_t o = word_shift(g, k);
    // loads at clamped (always valid) indices, then selects: no branch around the loads
    const uint32_t last = nlimbs - 1u;
    uint32_t lo = gmem(limbs)[k < last ? k : last];
    uint32_t hi = gmem(limbs)[k + 1u < last ? k + 1u : last];
    lo = k < nlimbs ? lo : 0u;
    hi = k + 1u < nlimbs ? hi : 0u;
    return (lo >> o) | (hi << (28u - o));
}

// A 32-bit input word as a residue: below P for P1 and P2, below 4P for P0 (the range of the
// lazy forward butterflies).  2^32 < 9.2 P0 < 2.2 P1 < 1.9 P2.  Each step subtracts when that
// does not wrap (x - m wraps above x exactly when x < m).
template <int I>
__device__ __forceinline__ uint32_t word_residue(uint32_t x) {
    if constexpr (I == 0) {
        constexpr uint32_t p4 = 4u * kP[0];
        uint32_t y = x - p4;
        x = y < x ? y : x;  // < 5.2 P0
        y = x - p4;
        return y < x ? y : x;  // < 4 P0
    } else if constexpr (lazy_prime(I)) {
        constexpr uint32_t p2 = 2u * kP[I];
        uint32_t y = x - p2;
        x = y < x ? y : x;  // < 2 P
        y = x - kP[I];
        return y < x ? y : x;
    } else {
        const uint32_t y = x - kP[I];  // x < 2^32 < 2 P
        return y < x ? y : x;
    }
}

template <int I>
__device__ __forceinline__ void words_to_residues(uint32_t (&x)[kE]) {
#pragma unroll
    for (int k = 0; k < kE; ++k) x[k] = word_residue<I>(x[k]);
}

// Where a forward transform's input comes from when it is read from limbs: words
// [off, off + cnt) of the operand (a chunk, or the whole operand with cnt = ~0u).
struct LimbSrc {
    const uint32_t* limbs;  // null: read the residue plane instead
    uint32_t nlimbs, off, cnt;
};

// Input words of a forward transform cut from limbs, in a group layout: element idx of the tile
// (row r, column c: idx = tile_base + c + (r << log_s)) is word src.off + idx of the operand, zero
// at and past src.cnt.  HALF: only the tile's lower half rows are read (in group 0's layout they
// are the elements q < EG / 2); the upper half is known to be zero.
template <int LR, int LC, int ST, int TL, bool HALF>
__device__ __forceinline__ void load_words(const LimbSrc& src, uint32_t (&x)[kE], uint32_t tile_base,
                                           uint32_t log_s) {
    using Ly = Layout<LR, LC, ST, TL>;
    constexpr int QN = HALF ? Ly::EG / 2 : Ly::EG;
    // Every limb load is issued before any word is cut, so the thread waits for memory once
    // (with the branch inside the element loop each word waited for its own two loads).
    // Limb indices are below 2^32; the loads index with them as 64-bit offsets.
    uint32_t lo[kE], hi[kE], g[kE];
    const uint32_t last = src.nlimbs - 1u;
#pragma unroll
    for (int sg = 0; sg < Ly::NSG; ++sg) {
        const uint32_t c = Ly::col(sg);
#pragma unroll
        for (int q = 0; q < QN; ++q) {
            const int e = sg * Ly::EG + q;
            const uint32_t idx = tile_base + c + (Ly::row(sg, q) << log_s);
            g[e] = idx < src.cnt ? src.off + idx : ~0u;  // ~0u: past the source, word 0
            const uint32_t k = word_limb(g[e] == ~0u ? 0u : g[e]);
            const uint32_t k0 = k < last ? k : last, k1 = k + 1u < last ? k + 1u : last;
            lo[e] = gmem(src.limbs)[(size_t)k0];
            hi[e] = gmem(src.limbs)[(size_t)k1];
        }
    }
#pragma unroll
    for (int sg = 0; sg < Ly::NSG; ++sg) {
#pragma unroll
        for (int q = 0; q < Ly::EG; ++q) {
            const int e = sg * Ly::EG + q;
            if (q >= QN) {
                x[e] = 0u;
                continue;
            }
            // bits [32g, 32g + 32) lie in limbs k = floor(32g / 28) and k + 1 from bit
            // o = 32g - 28k of limb k; limbs at and past nlimbs read as zero
            const uint32_t gg = g[e] == ~0u ? 0u : g[e];
            const uint32_t k = word_limb(gg), o = word_shift(gg, k);
            const uint32_t l = k < src.nlimbs ? lo[e] : 0u, h = k + 1u < src.nlimbs ? hi[e] : 0u;
            x[e] = g[e] == ~0u ? 0u : ((l >> o) | (h << (28u - o)));
        }
    }
}

// Words of the operand that can be nonzero: the source's words past its offset, at most cnt.
__device__ __forceinline__ uint32_t source_words(const LimbSrc& src) {
    const uint32_t w = words_of(src.nlimbs);
    const uint32_t avail = w > src.off ? w - src.off : 0u;
    return avail < src.cnt ? avail : src.cnt;
}

// Global memory <-> registers in a group layout.  Element (r, c) of the tile is at
// tile_base + c + (r << log_s) of the plane; with a limb source, element idx is input word
// src.off + idx (zero at and past src.cnt).
template <int LR, int LC, int ST, int TL, bool STORE>
__device__ __forceinline__ void global_io(uint32_t* __restrict__ plane, const LimbSrc& src,
                                          uint32_t (&x)[kE], uint32_t tile_base, uint32_t log_s) {
    using Ly = Layout<LR, LC, ST, TL>;
    if (!STORE && src.limbs) {
        load_words<LR, LC, ST, TL, false>(src, x, tile_base, log_s);
        return;
    }
#pragma unroll
    for (int sg = 0; sg < Ly::NSG; ++sg) {
        const uint32_t c = Ly::col(sg);
#pragma unroll
        for (int q = 0; q < Ly::EG; ++q) {
            const uint32_t idx = tile_base + c + (Ly::row(sg, q) << log_s);
            if (STORE) gmem(plane)[idx] = x[sg * Ly::EG + q];
            else x[sg * Ly::EG + q] = gmem(plane)[idx];
        }
    }
}

// Whole tile transform.  Forward: x arrives in group 0's layout and leaves in the last group's;
// inverse: the reverse (and scales by 2^LR).  SKIP0 / skip0: the tile's upper half rows are zero
// (see group_stages).
template <int I, bool INV, int LR, int LC, bool SKIP0 = false>
__device__ __forceinline__ void tile_transform(uint32_t (&x)[kE], uint32_t* lds,
                                               const uint32_t* tw, bool skip0 = false) {
    using S = Sched<LR>;
    static_for<S::NG>([&](auto step) {
        constexpr int gi = INV ? (S::NG - 1 - (int)step) : (int)step;
        if constexpr ((int)step > 0) {
            constexpr int pg = INV ? gi + 1 : gi - 1;
            __syncthreads();
            lds_io<LR, LC, S::stages(pg), S::trlo(pg), true>(lds, x);
            __syncthreads();
            lds_io<LR, LC, S::stages(gi), S::trlo(gi), false>(lds, x);
        }
        group_stages<I, INV, LR, LC, S::stages(gi), S::trlo(gi), SKIP0 && !INV && (int)step == 0>(x, tw, skip0);
    });
}

// ---------------------------------------------------------------------------------------------
// Outer (strided) pass: tile = 2^LR rows x 2^(12-LR) columns at row stride S = 2^log_s.
// grid.x = tiles, grid.y = job, grid.z = prime * nops + operand.
template <int I, bool INV, int LR>
__device__ __forceinline__ void outer_body(const PassArgs& a, const PassJob& job, uint32_t op,
                                           uint32_t* lds, uint32_t* tw) {
    constexpr int LC = 12 - LR;
    using S = Sched<LR>;
    const uint32_t n = 1u << a.log_n;
    const uint32_t ntiles = gridDim.x;
    uint32_t t = blockIdx.x;
    if ((ntiles & 7u) == 0u) t = (t & 7u) * (ntiles >> 3) + (t >> 3);  // neighbours on one XCD
    const uint32_t log_ncg = a.log_s - LC;
    const uint32_t blk = t >> log_ncg;
    const uint32_t cg = t & ((1u << log_ncg) - 1u);
    const uint32_t tile_base = (blk << (LR + a.log_s)) + (cg << LC);

    // both directions use the forward roots (see coeff_pos); the table's loads go out ahead of
    // the tile's
    TileTwiddles<I, LR, kThreads> tt;
    tt.load(a.wtab + I * kTile, block_pows<I>(a.beta, a.beta_stride, blk));

    // residue planes in and out: a first pass, which reads limbs, runs as k_outer_first
    uint32_t* plane = (op == 0 ? job.res_a : job.res_b) + (size_t)I * n;
    constexpr LimbSrc kPlane{nullptr, 0u, 0u, 0u};
    constexpr int G0 = 0, GL = S::NG - 1;
    uint32_t x[kE];
    if (!INV) {
        global_io<LR, LC, S::stages(G0), S::trlo(G0), false>(plane, kPlane, x, tile_base, a.log_s);
        tt.store(tw);
        __syncthreads();
        tile_transform<I, false, LR, LC>(x, lds, tw);
        global_io<LR, LC, S::stages(GL), S::trlo(GL), true>(plane, kPlane, x, tile_base, a.log_s);
    } else {
        global_io<LR, LC, S::stages(GL), S::trlo(GL), false>(plane, kPlane, x, tile_base, a.log_s);
        tt.store(tw);
        __syncthreads();
        tile_transform<I, true, LR, LC>(x, lds, tw);
        global_io<LR, LC, S::stages(G0), S::trlo(G0), true>(plane, kPlane, x, tile_base, a.log_s);
    }
}

// First forward pass, reading limbs: one block transforms its tile for all three primes, so the
// 32-bit words are cut from the limbs once and reduced per prime.  grid.z = operand.
template <int LR>
__device__ __forceinline__ void outer_first_body(const PassArgs& a, const PassJob& job, uint32_t op,
                                                 uint32_t* lds, uint32_t* tw) {
    constexpr int LC = 12 - LR;
    using S = Sched<LR>;
    const uint32_t n = 1u << a.log_n;
    const uint32_t ntiles = gridDim.x;
    uint32_t t = blockIdx.x;
    if ((ntiles & 7u) == 0u) t = (t & 7u) * (ntiles >> 3) + (t >> 3);  // neighbours on one XCD
    const uint32_t log_ncg = a.log_s - LC;
    const uint32_t blk = t >> log_ncg;
    const uint32_t cg = t & ((1u << log_ncg) - 1u);
    const uint32_t tile_base = (blk << (LR + a.log_s)) + (cg << LC);
    const LimbSrc src = op == 0 ? LimbSrc{job.limbs_a, job.la, job.a_off, job.a_cnt}
                                : LimbSrc{job.limbs_b, job.lb, 0u, ~0u};
    constexpr LimbSrc kPlane{nullptr, 0u, 0u, 0u};
    constexpr int G0 = 0, GL = S::NG - 1;
    // The first stage pairs rows r and r + R/2 (elements idx and idx + n/2).  When every upper
    // row of the tile lies past the operand (an operand filling at most half the transform, the
    // usual case), those rows are not read and the first stage is a copy.
    const bool half = tile_base + (1u << (LR - 1 + a.log_s)) >= source_words(src);
    // The three primes' twiddle tables (3 x 2^LR words) are built at once, their loads issued
    // ahead of the limb loads, so no prime waits for its table behind a barrier.
    TileTwiddles<0, LR, kThreads> tt0;
    TileTwiddles<1, LR, kThreads> tt1;
    TileTwiddles<2, LR, kThreads> tt2;
    tt0.load(a.wtab + 0 * kTile, block_pows<0>(a.beta, a.beta_stride, blk));
    tt1.load(a.wtab + 1 * kTile, block_pows<1>(a.beta, a.beta_stride, blk));
    tt2.load(a.wtab + 2 * kTile, block_pows<2>(a.beta, a.beta_stride, blk));
    uint32_t words[kE];
    if (half) load_words<LR, LC, S::stages(G0), S::trlo(G0), true>(src, words, tile_base, a.log_s);
    else load_words<LR, LC, S::stages(G0), S::trlo(G0), false>(src, words, tile_base, a.log_s);
    tt0.store(tw + (0 << LR));
    tt1.store(tw + (1 << LR));
    tt2.store(tw + (2 << LR));
    __syncthreads();
    static_for<kNumPrimes>([&](auto pi) {
        constexpr int I = (int)pi;
        // lds is reused across primes: the barrier in front of each exchange write orders it
        // after the previous prime's reads
        uint32_t x[kE];
#pragma unroll
        for (int k = 0; k < kE; ++k) x[k] = word_residue<I>(words[k]);
        tile_transform<I, false, LR, LC, true>(x, lds, tw + (I << LR), half);
        uint32_t* plane = (op == 0 ? job.res_a : job.res_b) + (size_t)I * n;
        global_io<LR, LC, S::stages(GL), S::trlo(GL), true>(plane, kPlane, x, tile_base, a.log_s);
    });
}

// Zero the launch sequence's carry-status words (PassArgs::clear) with every thread of the grid:
// the first kernel of a sequence does it, long before the sequence's Garner launch reads them.
__device__ __forceinline__ void clear_status(const PassArgs& a) {
    if (!a.clear) return;
    const uint32_t nthreads = blockDim.x * gridDim.x * gridDim.y * gridDim.z;
    const uint32_t id = ((blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
    for (uint32_t i = id; i < a.clear_words; i += nthreads) gmem(a.clear)[i] = 0u;
}

// Its own kernel, so the register demand of the three-prime loop does not lower the occupancy of
// the other strided passes.
template <int LR>
__global__ __launch_bounds__(kThreads) void k_outer_first(PassArgs a) {
    __shared__ uint32_t lds[kLdsWords];
    __shared__ uint32_t tw[kNumPrimes << LR];
    const PassJob job = a.jobs ? load_desc(a.jobs + blockIdx.y) : a.job;
    clear_status(a);
    outer_first_body<LR>(a, job, blockIdx.z, lds, tw);
}

template <bool INV, int LR>
__global__ __launch_bounds__(kThreads) void k_outer_pass(PassArgs a) {
    __shared__ uint32_t lds[kLdsWords];
    __shared__ uint32_t tw[1 << LR];
    const PassJob job = a.jobs ? load_desc(a.jobs + blockIdx.y) : a.job;
    clear_status(a);
    const uint32_t nops = INV ? 1u : ((a.flags & (kSquare | kSingle)) ? 1u : 2u);
    const uint32_t prime = blockIdx.z / nops, op = blockIdx.z % nops;
    switch (prime) {
        case 0: outer_body<0, INV, LR>(a, job, op, lds, tw); break;
        case 1: outer_body<1, INV, LR>(a, job, op, lds, tw); break;
        default: outer_body<2, INV, LR>(a, job, op, lds, tw); break;
    }
}

// ---------------------------------------------------------------------------------------------
// Block pass with the pointwise product: contiguous blocks of 2^LR points (one column).
// Forward-transform the block of A (and of B unless squaring), multiply pointwise, inverse
// transform, store into A's plane.  grid.x = n >> LR, grid.y = job, grid.z = prime.
// Pointwise product of forward outputs; P0 values arrive in [0, 4P) and leave in [0, 2P), the
// input range of the lazy inverse butterflies.
template <int I>
__device__ __forceinline__ uint32_t pointwise(uint32_t a, uint32_t b) {
    if constexpr (I == 0) return mont_mul_lazy<0>(fold2p<0>(a), fold2p<0>(b));
    else return mont_mul<I>(a, b);
}

// LIMBS: the pass reads the operands' limbs (a transform of at most one tile, no strided
// passes); its own instantiation keeps the limb path's registers out of the plane-reading one.
template <int I, int LR, bool LIMBS>
__device__ __forceinline__ void mid_body(const PassArgs& a, const PassJob& job, uint32_t* lds,
                                         uint32_t* tw) {
    using S = Sched<LR>;
    constexpr uint32_t T = 1u << (LR - 4);
    const uint32_t n = 1u << a.log_n;
    const uint32_t blk = blockIdx.x;
    const uint32_t tile_base = blk << LR;
    // one twiddle table per tile serves the forward and the inverse stages (see coeff_pos)
    TileTwiddles<I, LR, T> tt;
    tt.load(a.wtab + I * kTile, block_pows<I>(a.beta, a.beta_stride, blk));
    if (a.flags & kInvOnly) {  // a product spectrum from the diagonal sums: inverse stages only
        using Sx = Sched<LR>;
        constexpr LimbSrc kPl{nullptr, 0u, 0u, 0u};
        uint32_t* p = job.res_a + (size_t)I * n;
        uint32_t x[kE];
        global_io<LR, 0, Sx::stages(Sx::NG - 1), Sx::trlo(Sx::NG - 1), false>(p, kPl, x, tile_base, 0);
        tt.store(tw);
        __syncthreads();
        tile_transform<I, true, LR, 0>(x, lds, tw);
        global_io<LR, 0, Sx::stages(0), Sx::trlo(0), true>(p, kPl, x, tile_base, 0);
        return;
    }

    constexpr bool from_limbs = LIMBS;
    const bool square = (a.flags & kSquare) != 0;
    uint32_t* pa = job.res_a + (size_t)I * n;
    constexpr int G0 = 0;
    constexpr int GL = S::NG - 1;
    uint32_t xa[kE];
    constexpr LimbSrc kPlane{nullptr, 0u, 0u, 0u};
    const LimbSrc src_a = from_limbs ? LimbSrc{job.limbs_a, job.la, job.a_off, job.a_cnt} : kPlane;
    global_io<LR, 0, S::stages(G0), S::trlo(G0), false>(pa, src_a, xa, tile_base, 0);
    // B's residues are loaded before A is transformed where that costs no occupancy (2^11 and
    // 2^12-point tiles hold 16 more registers within the same waves per SIMD), after it otherwise
    constexpr bool kEarlyB = LR >= 11;
    const bool need_b = !square && !(a.flags & kFwdOnly);
    const bool bpre = (a.flags & kBPre) != 0;
    uint32_t* pb = job.res_b + (size_t)I * n;
    const LimbSrc src_b = from_limbs ? LimbSrc{job.limbs_b, job.lb, 0u, ~0u} : kPlane;
    uint32_t xb[kE];
    if (kEarlyB && need_b) {
        if (bpre) global_io<LR, 0, S::stages(GL), S::trlo(GL), false>(pb, kPlane, xb, tile_base, 0);
        else global_io<LR, 0, S::stages(G0), S::trlo(G0), false>(pb, src_b, xb, tile_base, 0);
    }
    tt.store(tw);
    if (from_limbs) words_to_residues<I>(xa);
    __syncthreads();  // forward twiddles ready
    tile_transform<I, false, LR, 0>(xa, lds, tw);
    if (a.flags & kFwdOnly) {  // store the transform in the last group's layout
        global_io<LR, 0, S::stages(GL), S::trlo(GL), true>(pa, kPlane, xa, tile_base, 0);
        return;
    }
    if (!square) {
        if (bpre) {
            if (!kEarlyB) global_io<LR, 0, S::stages(GL), S::trlo(GL), false>(pb, kPlane, xb, tile_base, 0);
        } else {
            if (!kEarlyB) global_io<LR, 0, S::stages(G0), S::trlo(G0), false>(pb, src_b, xb, tile_base, 0);
            if (from_limbs) words_to_residues<I>(xb);
            tile_transform<I, false, LR, 0>(xb, lds, tw);
        }
#pragma unroll
        for (int k = 0; k < kE; ++k) xa[k] = pointwise<I>(xa[k], xb[k]);
    } else {
#pragma unroll
        for (int k = 0; k < kE; ++k) xa[k] = pointwise<I>(xa[k], xa[k]);
    }
    tile_transform<I, true, LR, 0>(xa, lds, tw);
    global_io<LR, 0, S::stages(G0), S::trlo(G0), true>(pa, kPlane, xa, tile_base, 0);
}

// Block pass of a plain product or square (flags 0 or kSquare) over two adjacent tiles per
// block: the second tile's A residues are loaded before the first tile is computed and its B
// residues into the first tile's B registers once the pointwise product has consumed them, so
// both loads are in flight while the first tile is transformed.
template <int I, int LR>
__device__ __forceinline__ void mid_pair_body(const PassArgs& a, const PassJob& job, uint32_t* lds,
                                              uint32_t* tw) {
    using S = Sched<LR>;
    constexpr uint32_t T = 1u << (LR - 4);
    constexpr int G0 = 0;
    constexpr LimbSrc kPlane{nullptr, 0u, 0u, 0u};
    const uint32_t n = 1u << a.log_n;
    const bool square = (a.flags & kSquare) != 0;
    uint32_t* pa = job.res_a + (size_t)I * n;
    uint32_t* pb = job.res_b + (size_t)I * n;
    const uint32_t* wf = a.wtab + I * kTile;  // forward roots serve both directions
    const uint32_t blk0 = 2u * blockIdx.x;
    // Each tile's twiddle loads are issued ahead of the loads and stores in front of its table
    // build, so the build waits only for them.  The loop runs as a loop (its body is large), so the
    // reload for the second tile is unconditional: a conditional one would keep the first tile's
    // values live through the whole body.  On the second pass it reloads the second tile's.
    TileTwiddles<I, LR, T> tt;
    tt.load(wf, block_pows<I>(a.beta, a.beta_stride, blk0));
    uint32_t xa[kE], xb[kE], na[kE];
    global_io<LR, 0, S::stages(G0), S::trlo(G0), false>(pa, kPlane, xa, blk0 << LR, 0);
    if (!square) global_io<LR, 0, S::stages(G0), S::trlo(G0), false>(pb, kPlane, xb, blk0 << LR, 0);
    global_io<LR, 0, S::stages(G0), S::trlo(G0), false>(pa, kPlane, na, (blk0 + 1u) << LR, 0);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const uint32_t blk = blk0 + (uint32_t)t;
        if (t == 1) {
#pragma unroll
            for (int k = 0; k < kE; ++k) xa[k] = na[k];
            __syncthreads();  // the first tile is done with tw and lds
        }
        tt.store(tw);
        __syncthreads();
        tile_transform<I, false, LR, 0>(xa, lds, tw);
        if (!square) {
            tile_transform<I, false, LR, 0>(xb, lds, tw);
#pragma unroll
            for (int k = 0; k < kE; ++k) xa[k] = pointwise<I>(xa[k], xb[k]);
            if (t == 0)  // the second tile's B, into the registers just freed
                global_io<LR, 0, S::stages(G0), S::trlo(G0), false>(pb, kPlane, xb, (blk0 + 1u) << LR, 0);
        } else {
#pragma unroll
            for (int k = 0; k < kE; ++k) xa[k] = pointwise<I>(xa[k], xa[k]);
        }
        tile_transform<I, true, LR, 0>(xa, lds, tw);
        tt.load(wf, block_pows<I>(a.beta, a.beta_stride, blk0 + 1u));
        global_io<LR, 0, S::stages(G0), S::trlo(G0), true>(pa, kPlane, xa, blk << LR, 0);
    }
}

// grid.x = n >> (LR + 1) tile pairs, grid.y = job, grid.z = prime.
template <int LR>
__global__ __launch_bounds__(1 << (LR - 4)) void k_mid_pair(PassArgs a) {
    __shared__ uint32_t lds[(1 << LR) + (1 << LR) / 16];
    __shared__ uint32_t tw[1 << LR];
    const PassJob job = a.jobs ? load_desc(a.jobs + blockIdx.y) : a.job;
    clear_status(a);
    switch (blockIdx.z) {
        case 0: mid_pair_body<0, LR>(a, job, lds, tw); break;
        case 1: mid_pair_body<1, LR>(a, job, lds, tw); break;
        default: mid_pair_body<2, LR>(a, job, lds, tw); break;
    }
}

template <int LR, bool LIMBS>
__global__ __launch_bounds__(1 << (LR - 4)) void k_mid_pass(PassArgs a) {
    __shared__ uint32_t lds[(1 << LR) + (1 << LR) / 16];
    __shared__ uint32_t tw[1 << LR];
    const PassJob job = a.jobs ? load_desc(a.jobs + blockIdx.y) : a.job;
    clear_status(a);
    switch (blockIdx.z) {
        case 0: mid_body<0, LR, LIMBS>(a, job, lds, tw); break;
        case 1: mid_body<1, LR, LIMBS>(a, job, lds, tw); break;
        default: mid_body<2, LR, LIMBS>(a, job, lds, tw); break;
    }
}

// ---------------------------------------------------------------------------------------------
// Garner CRT + carry propagation (tile-local carry resolution and a fix-up pass).
//
// Coefficient j of the convolution of the operands' 32-bit words is recovered from the three
// inverse-transform outputs y_p (which carry a factor n * R^-1) as x_j < 2^90 (at most 2^26 terms
// below 2^64 each; P0 P1 P2 > 2^90.8) and split into 32-bit pieces w0, w1, w2.  Output word i of
// the product is w0_i + w1_{i-1} + w2_{i-2} + carry_i: each sum is below 3 * 2^32, so carry_i is
// in {0, 1, 2} and the carry-in -> carry-out behaviour of any run of words is a map
// {0..3} -> {0..3}, stored as four 2-bit fields.  Maps compose associatively, which makes the
// carry chain a scan.  The product's words are finally cut back into 28-bit limbs: a thread owns
// 7 words = 8 limbs, a tile 3584 words = 4096 limbs, so limbs never straddle threads.

__device__ __forceinline__ uint32_t map_apply(uint32_t m, uint32_t c) { return (m >> (2u * c)) & 3u; }
// (second after first): apply `first`, then `second`
__device__ __forceinline__ uint32_t map_compose(uint32_t second, uint32_t first) {
    uint32_t r = 0;
#pragma unroll
    for (uint32_t c = 0; c < 4; ++c) r |= map_apply(second, map_apply(first, c)) << (2u * c);
    return r;
}
constexpr uint32_t kMapIdentity = 0b11100100u;  // c -> c
constexpr uint32_t kFixBit = 1u << 8;           // status word: tile written with carry-in 0

struct Pieces {
    uint32_t w0, w1, w2;  // x = w0 + w1 2^32 + w2 2^64
};

// [0, 2 P0) -> [0, P0)
__device__ __forceinline__ uint32_t fold1p(uint32_t x) {
    const uint32_t y = x - kP[0];
    return y < x ? y : x;
}

// Garner constants of one job as the CRT uses them.  ks[i] = n^-1 R^2 mod P_i is the job's scale
// (mont_mul(y, ks[i]) = y n^-1 R); a1 = ks[1] inv01 R^-1 and a2 = ks[2] inv02 R^-1 fold it into the
// first Garner products, and m01 = P1 - inv01, m02 = P2 - inv02 are their negated factors.
struct CrtConsts {
    uint32_t ks0, a1, a2, m01, m02, inv12;
};

__device__ __forceinline__ CrtConsts crt_consts(const uint32_t (&ks)[3], const GarnerConsts& k) {
    return CrtConsts{ks[0], mont_mul<1>(ks[1], k.inv01), mont_mul<2>(ks[2], k.inv02), kP[1] - k.inv01,
                     kP[2] - k.inv02, k.inv12};
}

// Coefficient from its three inverse-transform outputs r0, r1, r2 (each still carrying the
// factor n * R^-1); r0 may lie in [0, 2 P0), r1 < P1, r2 < P2.  Garner's
//   v0 = c0,  v1 = (c1 - v0) inv01 mod P1,  t = (c2 - v0) inv02 mod P2,  v2 = (t - v1) inv12 mod P2
// with c_i = r_i n^-1 (mod P_i) computes v1 and t each as one Montgomery reduction of a two-product
// sum: r1 a1 + v0 (P1 - inv01) < P1^2 + P0 P1 < 4.1e18 + 9.5e17 and r2 a2 + v0 (P2 - inv02)
// < P2^2 + P0 P2 < 5.3e18 + 1.1e18, both inside mont_redc's bound because v0 < P0 is small.
__device__ __forceinline__ unsigned __int128 crt_coeff(uint32_t r0, uint32_t r1, uint32_t r2,
                                                       const CrtConsts& c) {
    const uint32_t v0 = mont_mul<0>(r0, c.ks0);
    // v0 < P0 < P1 < P2 and v1 < P1 < P2 are already residues of the larger primes
    static_assert(kP[0] < kP[1] && kP[1] < kP[2], "CRT order");
    const uint32_t v1 = mont_redc<1>((uint64_t)r1 * c.a1 + (uint64_t)v0 * c.m01);
    const uint32_t t = mont_redc<2>((uint64_t)r2 * c.a2 + (uint64_t)v0 * c.m02);
    const uint32_t v2 = mont_mul<2>(sub_mod<2>(t, v1), c.inv12);
    const uint64_t inner = (uint64_t)v1 + (uint64_t)kP[1] * v2;  // < 2^62.1
    return (unsigned __int128)inner * kP[0] + v0;
}

// Position of convolution coefficient j (j < n) in an inverse-transform output.  The inverse
// runs the Gentleman-Sande stages with the forward roots (the forward twiddle tables serve both
// directions), which computes the transform with root w instead of w^-1: its output holds
// n * x[(n - j) mod n] at position j, so coefficient j sits at position (n - j) mod n.
__device__ __forceinline__ uint32_t coeff_pos(uint32_t j, uint32_t n) { return (n - j) & (n - 1u); }

// Words wx and wx - 1 (wx even) of one plane of n inverse-transform outputs, words outside
// [0, n) reading as zero.  Word w sits at position (n - w) mod n, so for 2 <= wx <= n - 2 the two
// words are the aligned position pair (n - wx, n - wx + 1); word 0 is at position 0 and word n - 1
// at position 1.  Every wx does one aligned load inside the plane (the pair at position 0 where
// wx is not in [2, n - 2]) and selects the words it needs, with no branch around the load, so a
// caller's loads issue back to back and are waited for together; a branch per load makes the
// compiler wait for each load before it issues the next.
__device__ __forceinline__ u32x2 load_pair_rev(const uint32_t* __restrict__ yp, uint32_t n, int64_t wx) {
    const bool mid = wx >= 2 && wx <= (int64_t)n - 2;
    const u32x2 v = *gmem(reinterpret_cast<const u32x2*>(yp + (mid ? n - (uint32_t)wx : 0u)));
    return u32x2{mid || wx == 0 ? v.x : 0u, mid || wx == (int64_t)n ? v.y : 0u};
}

// The coefficient that wrapped from index n onto index 0: product of the operands' top words.
__device__ __forceinline__ uint64_t wrap_product(const GarnerJob& jb) {
    const uint32_t ga = words_of(jb.wrap_la) - 1u, gb = words_of(jb.wrap_lb) - 1u;
    return (uint64_t)limb_word(jb.wrap_a, jb.wrap_la, ga) * limb_word(jb.wrap_b, jb.wrap_lb, gb);
}

// acc + x mod P_p for a residue x of prime p's inverse output (P0 outputs lie in [0, 2 P0)).
__device__ __forceinline__ uint32_t add_residue(uint32_t p, uint32_t acc, uint32_t x) {
    return p == 0u ? add_mod<0>(acc, fold1p(x)) : (p == 1u ? add_mod<1>(acc, x) : add_mod<2>(acc, x));
}

constexpr uint32_t kGThreads = 512;           // threads per Garner + carry tile
constexpr uint32_t kCarryWordsPerThread = 7;   // 7 * 32 = 8 * 28 bits
constexpr uint32_t kCarryLimbsPerThread = 8;
constexpr uint32_t kCarryTile = kGThreads * kCarryLimbsPerThread;  // 4096 limbs per tile
constexpr uint32_t kWordTile = kGThreads * kCarryWordsPerThread;   // 3584 words per tile
static_assert(kCarryTile == kCarryTileLimbs && kWordTile == kCarryTileWords, "tile geometry");

// Coefficient j of a job from its three (chunk-summed) residues, split into 32-bit pieces.  An
// unchunked product whose top coefficient wrapped from index n onto index 0 has it taken out of
// coefficient 0 and put back at index n (whose residues read as zero); `wrap` is false where the
// caller knows j is neither (a tile away from both ends).
__device__ __forceinline__ Pieces coeff_pieces(uint32_t r0, uint32_t r1, uint32_t r2, int64_t j,
                                               const GarnerJob& jb, const CrtConsts& c, bool wrap) {
    unsigned __int128 x = crt_coeff(r0, r1, r2, c);
    if (wrap) {
        if (j == 0) x -= wrap_product(jb);
        else if (j == (int64_t)jb.n) x = wrap_product(jb);
    }
    const uint64_t lo = (uint64_t)x;
    return Pieces{(uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)(x >> 64)};
}


// CHUNKED: the jobs are chunked products (nchunks > 1); otherwise every job is unchunked.  Two
// instantiations keep the chunk loop's register demand out of the common kernel.
//
// A tile of 3584 words (4096 limbs) per block of 512 threads; thread t owns words
// j0 = base + 7t .. j0 + 6.  Its word sums s_q = w0(j0+q) + w1(j0+q-1) + w2(j0+q-2) take the
// pieces of coefficients j0 - 1 and j0 - 2 from its left neighbour (a wave shuffle, across waves
// through LDS); thread 0 takes them from coefficients base - 1 and base - 2, which threads 0..2
// reconstruct together with base - 3 (for the tile's local carry rule).
template <bool CHUNKED>
__device__ __forceinline__ void garner_tile(const GarnerArgs& a, const GarnerJob& jb, uint32_t tile) {
    constexpr uint32_t kWaves = kGThreads / 64;
    __shared__ uint32_t s_wave[kWaves];     // slow path: each wave's inclusive carry map
    __shared__ uint32_t s_wc[kWaves];       // lane 63 of each wave: carry-out of its run for carry-in 0
    __shared__ uint32_t s_wconst[kWaves];   // each wave: all its runs pass on a fixed carry
    __shared__ uint32_t s_carry_in;
    __shared__ uint32_t s_edge[kWaves][3];  // lane 63 of each wave: w1 and w2 of its last word, w2 of the one before
    __shared__ uint32_t s_halo[3][3];       // pieces of words base - 3, base - 2, base - 1
    const uint32_t tid = threadIdx.x;
    const uint32_t base = tile * kWordTile;  // first word (coefficient) of the tile
    const uint32_t n = jb.n;
    // An interior tile reads only aligned position pairs inside the plane (its window
    // [base - 4, base + 3584] lies in [2, n - 2]) and none of its coefficients is index 0 or n.
    const bool interior = !CHUNKED && base >= 6u && (uint64_t)base + kWordTile + 2u <= n;
    const CrtConsts crt = crt_consts(jb.kscale, a.k);
    const bool wrap = jb.wrap_a != nullptr && !interior;

    // 1. Residues.  The block's words [base - 4, base + kWordTile) of each plane are read coalesced
    //    and staged in LDS one plane at a time, where each thread picks up its run of 7.  Word w
    //    sits at position (n - w) mod n (coeff_pos), so aligned position pairs hold descending
    //    word pairs: pair i of the window is words base + 2i and base + 2i - 1.  Thread tid loads
    //    pairs i = k * 512 + tid (k < 4, i < 1792: words base - 1 .. base + 3582), threads 0, 1, 2
    //    the window's ends (pairs -2, -1: words base - 4 .. base - 2; pair 1792: word base + 3583).
    //    Words outside every transform read as zero; a chunked product's chunk residues are summed
    //    on the way.  The output limbs go through the same LDS array.
    constexpr uint32_t kStage = kWordTile + 4u;  // 4 halo words + tile
    constexpr uint32_t kResWords = kStage > kCarryTile ? kStage : kCarryTile;
    __shared__ __align__(16) uint32_t s_res[kResWords];
    constexpr uint32_t kNPairs = kWordTile / 2u;                       // 1792
    constexpr uint32_t kPairLoads = (kNPairs + kGThreads - 1) / kGThreads;  // 4
    const int64_t ie = tid == 0u ? -2 : (tid == 1u ? -1 : (int64_t)kNPairs);
    const bool has_end = tid < 3u;
    uint32_t r[3][kCarryWordsPerThread];  // residues of the thread's words, per plane
    uint32_t hr[3] = {0u, 0u, 0u};        // threads 0..2: residues of word base - 3 + tid
    {
        u32x2 v[3][kPairLoads], h[3] = {{0u, 0u}, {0u, 0u}, {0u, 0u}};
#pragma unroll
        for (uint32_t p = 0; p < 3u; ++p)
#pragma unroll
            for (uint32_t k = 0; k < kPairLoads; ++k) v[p][k] = u32x2{0u, 0u};
        if (interior) {
            // pair i (words base + 2i, base + 2i - 1) at positions n - base - 2i, +1
#pragma unroll
            for (uint32_t p = 0; p < 3u; ++p) {
                const uint32_t* yp = jb.y + (size_t)p * n + (n - base);
#pragma unroll
                for (uint32_t k = 0; k < kPairLoads; ++k) {
                    // unconditional loads, all in the plane (lanes past the tile's pairs reload
                    // pair 0; ie = 1792 of the lanes without a halo pair is the tile's last pair),
                    // so they issue back to back; only the wanted values are staged
                    const uint32_t i = k * kGThreads + tid;
                    v[p][k] = *gmem(reinterpret_cast<const u32x2*>(yp - 2u * (i < kNPairs ? i : 0u)));
                }
                h[p] = *gmem(reinterpret_cast<const u32x2*>(yp - 2 * ie));
            }
        } else if constexpr (!CHUNKED) {
#pragma unroll
            for (uint32_t p = 0; p < 3u; ++p) {
                const uint32_t* yp = jb.y + (size_t)p * n;
#pragma unroll
                for (uint32_t k = 0; k < kPairLoads; ++k) {
                    // lanes past the tile's pairs read zeros (wx = -2), with no branch around the load
                    const uint32_t i = k * kGThreads + tid;
                    v[p][k] = load_pair_rev(yp, n, i < kNPairs ? (int64_t)base + 2 * (int64_t)i : -2);
                }
                h[p] = load_pair_rev(yp, n, has_end ? (int64_t)base + 2 * ie : -2);
            }
        } else {
            // chunk cc covers words [cc L, cc L + n); those reaching [lo, hi] are
            // (lo - n) / L < cc <= hi / L
            const uint32_t L = jb.chunk_len;
            const bool p2 = jb.log_chunk < 32u;
            const uint32_t lo = base >= 4u ? base - 4u : 0u, hi = base + kWordTile - 1u;
            uint32_t c_hi = p2 ? (hi >> jb.log_chunk) : hi / L;
            c_hi = c_hi < jb.nchunks - 1u ? c_hi : jb.nchunks - 1u;
            const uint32_t c_lo = lo >= n ? (p2 ? ((lo - n) >> jb.log_chunk) : (lo - n) / L) + 1u : 0u;
            for (uint32_t cc = c_lo; cc <= c_hi; ++cc) {
                const uint32_t* yc = jb.y + (size_t)cc * jb.chunk_stride;
                const int64_t c0 = (int64_t)cc * L;  // first word of chunk cc
                // the chunk's loads first, all in flight together, then the sums
                u32x2 x[3][kPairLoads], xh[3];
#pragma unroll
                for (uint32_t p = 0; p < 3u; ++p) {
                    const uint32_t* yp = yc + (size_t)p * n;
#pragma unroll
                    for (uint32_t k = 0; k < kPairLoads; ++k) {
                        const uint32_t i = k * kGThreads + tid;
                        x[p][k] = load_pair_rev(yp, n, i < kNPairs ? (int64_t)base + 2 * (int64_t)i - c0 : -2);
                    }
                    xh[p] = load_pair_rev(yp, n, has_end ? (int64_t)base + 2 * ie - c0 : -2);
                }
#pragma unroll
                for (uint32_t p = 0; p < 3u; ++p) {
#pragma unroll
                    for (uint32_t k = 0; k < kPairLoads; ++k)
                        v[p][k] = u32x2{add_residue(p, v[p][k].x, x[p][k].x), add_residue(p, v[p][k].y, x[p][k].y)};
                    h[p] = u32x2{add_residue(p, h[p].x, xh[p].x), add_residue(p, h[p].y, xh[p].y)};
                }
            }
        }
        const uint32_t o = tid * kCarryWordsPerThread + 4u;  // staged index of word j0
#pragma unroll
        for (uint32_t p = 0; p < 3u; ++p) {
            if (p > 0u) __syncthreads();  // the previous plane has been read
            // pair i: word base + 2i at staged index 2i + 4, word base + 2i - 1 at 2i + 3
#pragma unroll
            for (uint32_t k = 0; k < kPairLoads; ++k) {
                const uint32_t i = k * kGThreads + tid;
                if (i < kNPairs) {
                    s_res[2u * i + 3u] = v[p][k].y;
                    s_res[2u * i + 4u] = v[p][k].x;
                }
            }
            if (tid == 0u) s_res[0] = h[p].x;  // pair -2: word base - 4 (base - 5 is outside)
            if (tid == 1u) {                   // pair -1: words base - 2, base - 3
                s_res[1] = h[p].y;
                s_res[2] = h[p].x;
            }
            if (tid == 2u) s_res[kStage - 1u] = h[p].y;  // pair 1792: word base + 3583
            __syncthreads();
#pragma unroll
            for (uint32_t q = 0; q < kCarryWordsPerThread; ++q) r[p][q] = s_res[o + q];
            if (has_end) hr[p] = s_res[1u + tid];  // words base - 3, base - 2, base - 1
        }
    }

    // 2. Pieces and word sums.  s_q (64-bit: lo, hi) collects w0(j0+q) + w1(j0+q-1) + w2(j0+q-2)
    //    of the thread's own coefficients; the neighbour's share of s_0 and s_1 follows below.
    const int64_t j0 = (int64_t)base + (int64_t)tid * kCarryWordsPerThread;
    uint32_t s_lo[kCarryWordsPerThread], s_hi[kCarryWordsPerThread];
    uint32_t e_w1 = 0u, e_w2 = 0u, e_w2p = 0u;  // w1, w2 of word j0 + 6, w2 of word j0 + 5
    {
        uint64_t acc1 = 0, acc2 = 0;  // partial sums of the next word and the one after
#pragma unroll
        for (uint32_t q = 0; q < kCarryWordsPerThread; ++q) {
            const Pieces pc = coeff_pieces(r[0][q], r[1][q], r[2][q], j0 + q, jb, crt, wrap);
            const uint64_t sum = acc1 + pc.w0;
            s_lo[q] = (uint32_t)sum;
            s_hi[q] = (uint32_t)(sum >> 32);
            acc1 = acc2 + pc.w1;
            acc2 = pc.w2;
            if (q == kCarryWordsPerThread - 2u) e_w2p = pc.w2;
            if (q == kCarryWordsPerThread - 1u) {
                e_w1 = pc.w1;
                e_w2 = pc.w2;
            }
        }
    }
    if (has_end) {
        const Pieces pc = coeff_pieces(hr[0], hr[1], hr[2], (int64_t)base - 3 + (int64_t)tid, jb, crt, wrap);
        s_halo[tid][0] = pc.w0;
        s_halo[tid][1] = pc.w1;
        s_halo[tid][2] = pc.w2;
    }
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    if (lane == 63u) {
        s_edge[wave][0] = e_w1;
        s_edge[wave][1] = e_w2;
        s_edge[wave][2] = e_w2p;
    }
    // the left neighbour's w1(j0 - 1), w2(j0 - 1), w2(j0 - 2)
    uint32_t n_w1 = __shfl_up(e_w1, 1, 64), n_w2 = __shfl_up(e_w2, 1, 64), n_w2p = __shfl_up(e_w2p, 1, 64);
    __syncthreads();
    if (lane == 0u) {
        if (wave > 0u) {
            n_w1 = s_edge[wave - 1u][0];
            n_w2 = s_edge[wave - 1u][1];
            n_w2p = s_edge[wave - 1u][2];
        } else {
            n_w1 = s_halo[2][1];   // word base - 1
            n_w2 = s_halo[2][2];
            n_w2p = s_halo[1][2];  // word base - 2
        }
    }
    {
        // s_0 += w1(j0 - 1) + w2(j0 - 2), s_1 += w2(j0 - 1): below 3 * 2^32 in total
        uint64_t t0 = ((uint64_t)s_hi[0] << 32 | s_lo[0]) + n_w1 + n_w2p;
        uint64_t t1 = ((uint64_t)s_hi[1] << 32 | s_lo[1]) + n_w2;
        s_lo[0] = (uint32_t)t0;
        s_hi[0] = (uint32_t)(t0 >> 32);
        s_lo[1] = (uint32_t)t1;
        s_hi[1] = (uint32_t)(t1 >> 32);
    }

    // The carry into the tile, resolved locally when possible: the previous tile's last word has
    // the sum s = w0(base-1) + w1(base-2) + w2(base-3) and a carry-in of at most 2, so its
    // carry-out is s >> 32 whatever that carry-in, unless s mod 2^32 > 2^32 - 3.  Then the tile
    // needs nothing from its predecessors (0xFF: not resolved, left to k_carry_fixup).
    uint32_t local = 0xFFu;
    if (tid == 0 && tile != 0) {
        const uint64_t sl = (uint64_t)s_halo[2][0] + s_halo[1][1] + s_halo[0][2];
        if ((uint32_t)sl <= 0xFFFFFFFDu) local = (uint32_t)(sl >> 32);
    }

    // 3. Carries inside the tile.  The carry into a word is at most 2, and a run's carry-out is a
    //    non-decreasing function of its carry-in, so a run whose carry-outs for carry-ins 0 and 2
    //    agree passes on that carry whatever it receives (in random data every run does: only a
    //    word sum within 2 of a multiple of 2^32 can propagate).  When all runs of the tile are
    //    such, the carry into run t is the carry-out of run t - 1 and the tile's own carry-out is
    //    that of its last run.  Otherwise each run's carry map (carry-out for each carry-in 0..3,
    //    four chains side by side) is scanned across the block.
    uint32_t co0 = 0u, co2 = 2u;
#pragma unroll
    for (uint32_t q = 0; q < kCarryWordsPerThread; ++q) {
        const uint32_t v0 = s_lo[q] + co0;
        co0 = s_hi[q] + (v0 < co0 ? 1u : 0u);
        const uint32_t v2 = s_lo[q] + co2;
        co2 = s_hi[q] + (v2 < co2 ? 1u : 0u);
    }
    const bool wconst = __ballot(co0 == co2) == ~0ull;
    if (lane == 63u) {
        s_wc[wave] = co0;
        s_wconst[wave] = wconst ? 1u : 0u;
    }
    __syncthreads();
    bool fixed_runs = true;
#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) fixed_runs = fixed_runs && s_wconst[w] != 0u;
    uint32_t excl = kMapIdentity;  // slow path: composition of the maps of runs 0 .. t-1
    uint32_t agg;                  // the tile's carry map
    if (fixed_runs) {
        agg = s_wc[kWaves - 1u] * 0x55u;  // a constant map
    } else {
        uint32_t m;
        {
            uint32_t c[4] = {0u, 1u, 2u, 3u};
#pragma unroll
            for (uint32_t q = 0; q < kCarryWordsPerThread; ++q) {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t v = s_lo[q] + c[k];
                    c[k] = s_hi[q] + (v < c[k] ? 1u : 0u);
                }
            }
            m = c[0] | (c[1] << 2) | (c[2] << 4) | (c[3] << 6);
        }
        uint32_t inc = m;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(inc, d, 64);
            if (lane >= d) inc = map_compose(inc, o);
        }
        if (lane == 63) s_wave[wave] = inc;
        __syncthreads();
        uint32_t wave_prefix = kMapIdentity;
        for (uint32_t w = 0; w < wave; ++w) wave_prefix = map_compose(s_wave[w], wave_prefix);
        excl = __shfl_up(inc, 1, 64);
        if (lane == 0) excl = kMapIdentity;
        excl = map_compose(excl, wave_prefix);
        agg = kMapIdentity;
        for (uint32_t w = 0; w < kWaves; ++w) agg = map_compose(s_wave[w], agg);
    }

    // 4. The carry into the tile: 0 for tile 0, else resolved locally (above).  A tile that cannot
    //    resolve it is written as if it were 0 and flagged; k_carry_fixup then scans the tiles'
    //    carry maps and adds the true carry-in to such tiles.  No block waits for another, so the
    //    tiles need no ordering and no counter.  Every tile publishes its carry map.
    if (tid == 0) {
        const bool unresolved = tile != 0 && local == 0xFFu;
        gmem(jb.status)[tile] = (unresolved ? kFixBit : 0u) | agg;
        if (unresolved) gmem(jb.fix)[0] = 1u;
        s_carry_in = tile == 0 || unresolved ? 0u : local;
    }
    // the carry-out of run t - 1 (fast path), read before the barrier so every lane takes part
    uint32_t prev = __shfl_up(co0, 1, 64);
    if (lane == 0u) prev = wave > 0u ? s_wc[wave - 1u] : 0u;
    __syncthreads();

    // 5. apply the carries, cut the 7 words into 8 limbs and write them
    uint32_t carry = fixed_runs ? (tid == 0u ? s_carry_in : prev) : map_apply(excl, s_carry_in);
    uint32_t wv[kCarryWordsPerThread];
#pragma unroll
    for (uint32_t q = 0; q < kCarryWordsPerThread; ++q) {
        const uint32_t v = s_lo[q] + carry;
        carry = s_hi[q] + (v < carry ? 1u : 0u);
        wv[q] = v;
    }
    uint32_t outv[kCarryLimbsPerThread];
#pragma unroll
    for (uint32_t q = 0; q < kCarryLimbsPerThread; ++q) {
        const uint32_t bit = 28u * q, wi = bit >> 5, bo = bit & 31u;
        uint32_t v = wv[wi] >> bo;
        if (bo > 4u) v |= wv[wi + 1] << (32u - bo);
        outv[q] = v & kDigitMask;
    }
    // through LDS (every thread has read its residues: barriers passed since), then written
    // lane-contiguous
#pragma unroll
    for (uint32_t q = 0; q < kCarryLimbsPerThread / 4; ++q)
        *reinterpret_cast<u32x4*>(s_res + tid * kCarryLimbsPerThread + 4u * q) =
            u32x4{outv[4 * q], outv[4 * q + 1], outv[4 * q + 2], outv[4 * q + 3]};
    __syncthreads();
    uint32_t* out = jb.out + (size_t)tile * kCarryTile;
    const uint32_t lim = jb.out_len - tile * kCarryTile;  // limbs of this tile to write
    if ((reinterpret_cast<uintptr_t>(out) & 15u) == 0) {
#pragma unroll
        for (uint32_t k = 0; k < kCarryLimbsPerThread / 4; ++k) {
            const uint32_t idx = 4u * (k * kGThreads + tid);
            const u32x4 x = *reinterpret_cast<const u32x4*>(s_res + idx);
            if (idx + 4u <= lim) {
                *gmem(reinterpret_cast<u32x4*>(out + idx)) = x;
            } else {
                if (idx < lim) gmem(out)[idx] = x.x;
                if (idx + 1u < lim) gmem(out)[idx + 1u] = x.y;
                if (idx + 2u < lim) gmem(out)[idx + 2u] = x.z;
            }
        }
    } else {
#pragma unroll
        for (uint32_t m2 = 0; m2 < kCarryLimbsPerThread; ++m2) {
            const uint32_t idx = m2 * kGThreads + tid;
            if (idx < lim) gmem(out)[idx] = s_res[idx];
        }
    }
}

// One tile per block (grid.x = tiles, grid.y = job).
template <bool CHUNKED>
__global__ __launch_bounds__(kGThreads) void k_garner_carry(GarnerArgs a) {
    const GarnerJob jb = a.jobs ? load_desc(a.jobs + blockIdx.y) : a.job;
    const uint32_t ntiles = (jb.out_len + kCarryTile - 1u) / kCarryTile;
    if (blockIdx.x < ntiles) garner_tile<CHUNKED>(a, jb, blockIdx.x);
}

// Carry fix-up of one job (grid.y = job, one block of kFixThreads threads): nothing to do unless a
// tile of the job was written with an assumed carry-in of 0.  Then the block scans the tiles'
// carry maps in chunks of kFixThreads tiles (carry_in(0) = 0, carry_in(t + 1) = map_t(carry_in(t)))
// and adds each flagged tile's true carry-in c at its first limb: the tile's limbs hold its value
// with carry-in 0, and its value with carry-in c is that plus c modulo the tile (what passes out
// of the tile is the next tile's carry-in, which the maps account for).  The addition changes
// limb 0 and, when it overflows, clears the run of all-ones limbs after it and increments the
// first limb that is not all ones.
constexpr uint32_t kFixThreads = 1024;

__global__ __launch_bounds__(kFixThreads) void k_carry_fixup(GarnerArgs a) {
    const GarnerJob jb = a.jobs ? load_desc(a.jobs + blockIdx.y) : a.job;
    if (gmem(jb.fix)[0] == 0u) return;
    constexpr uint32_t kWaves = kFixThreads / 64;
    __shared__ uint32_t s_wmap[kWaves];
    __shared__ uint32_t s_carry;   // carry into the current chunk's first tile
    __shared__ uint32_t s_first;   // the ripple's first limb that is not all ones
    __shared__ uint32_t s_nfix;
    __shared__ uint32_t s_fix[kFixThreads][2];  // tiles of the chunk to fix: tile, carry-in
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t ntiles = (jb.out_len + kCarryTile - 1u) / kCarryTile;
    if (tid == 0) s_carry = 0u;
    for (uint32_t t0 = 0; t0 < ntiles; t0 += kFixThreads) {
        const uint32_t t = t0 + tid;
        const uint32_t st = t < ntiles ? gmem(jb.status)[t] : kMapIdentity;
        const uint32_t m = st & 0xFFu;
        // exclusive scan of the maps across the block: inclusive within the wave, then waves
        uint32_t inc = m;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(inc, d, 64);
            if (lane >= d) inc = map_compose(inc, o);
        }
        if (lane == 63u) s_wmap[wave] = inc;
        if (tid == 0) s_nfix = 0u;
        __syncthreads();
        uint32_t pre = kMapIdentity;
        for (uint32_t w = 0; w < wave; ++w) pre = map_compose(s_wmap[w], pre);
        uint32_t excl = __shfl_up(inc, 1, 64);
        if (lane == 0) excl = kMapIdentity;
        excl = map_compose(excl, pre);
        const uint32_t cin = map_apply(excl, s_carry);  // carry into tile t
        if (t < ntiles && (st & kFixBit) && cin != 0u) {
            const uint32_t k = atomicAdd(&s_nfix, 1u);
            s_fix[k][0] = t;
            s_fix[k][1] = cin;
        }
        uint32_t chunk = kMapIdentity;
        for (uint32_t w = 0; w < kWaves; ++w) chunk = map_compose(s_wmap[w], chunk);
        __syncthreads();  // s_carry, s_wmap and the fix list are complete
        const uint32_t carry_next = map_apply(chunk, s_carry);
        const uint32_t nfix = s_nfix;
        for (uint32_t f = 0; f < nfix; ++f) {
            const uint32_t tile = s_fix[f][0], c = s_fix[f][1];
            uint32_t* out = jb.out + (size_t)tile * kCarryTile;
            const uint32_t lim = min(kCarryTile, jb.out_len - tile * kCarryTile);
            if (tid == 0) s_first = lim;
            __syncthreads();
            const uint32_t v0 = gmem(out)[0] + c;
            if (v0 > kDigitMask) {
                // ripple: find the first limb after limb 0 that is not all ones
                for (uint32_t i0 = 1; i0 < lim; i0 += kFixThreads) {
                    const uint32_t i = i0 + tid;
                    if (i < lim && gmem(out)[i] != kDigitMask) atomicMin(&s_first, i);
                    __syncthreads();
                    const uint32_t found = s_first;
                    __syncthreads();  // every thread has read s_first before the next round lowers it
                    if (found < lim) break;
                }
                const uint32_t first = s_first;
                for (uint32_t i = 1 + tid; i < first; i += kFixThreads) gmem(out)[i] = 0u;
                if (tid == 0 && first < lim) gmem(out)[first] += 1u;
            }
            __syncthreads();  // every thread has read limb 0 (and s_first) before they change
            if (tid == 0) gmem(out)[0] = v0 & kDigitMask;
            __syncthreads();
        }
        if (tid == 0) s_carry = carry_next;
        __syncthreads();
    }
}

// Diagonal sums of chunk spectra (see launch_diag_accumulate).  grid.x covers a plane,
// grid.y = diagonal, grid.z = prime.  P0 spectra arrive in [0, 4P) from the lazy forward stages.
template <int I>
__device__ __forceinline__ void diag_body(const uint32_t* __restrict__ sa, const uint32_t* __restrict__ sb,
                                          uint32_t ca, uint32_t cb, uint32_t n, uint32_t* __restrict__ d,
                                          uint32_t s, uint32_t e) {
    const size_t stride = (size_t)3 * n, off = (size_t)I * n + e;
    const uint32_t i_lo = s + 1u > cb ? s + 1u - cb : 0u;
    const uint32_t i_hi = s < ca - 1u ? s : ca - 1u;
    uint32_t acc = 0;
    for (uint32_t i = i_lo; i <= i_hi; ++i) {
        const uint32_t x = gmem(sa)[i * stride + off], y = gmem(sb)[(s - i) * stride + off];
        if constexpr (I == 0) acc = add_mod<0>(acc, fold1p(pointwise<0>(x, y)));
        else acc = add_mod<I>(acc, mont_mul<I>(x, y));
    }
    gmem(d)[s * stride + off] = acc;
}

__global__ __launch_bounds__(256) void k_diag_accumulate(const uint32_t* sa, const uint32_t* sb, uint32_t ca,
                                                         uint32_t cb, uint32_t n, uint32_t* d) {
    const uint32_t e = blockIdx.x * 256u + threadIdx.x;
    if (e >= n) return;
    switch (blockIdx.z) {
        case 0: diag_body<0>(sa, sb, ca, cb, n, d, blockIdx.y, e); break;
        case 1: diag_body<1>(sa, sb, ca, cb, n, d, blockIdx.y, e); break;
        default: diag_body<2>(sa, sb, ca, cb, n, d, blockIdx.y, e); break;
    }
}


}  // namespace bi

// ---------------------------------------------------------------------------------------------
// Launch wrappers (host side of this translation unit).
namespace bi {

using KernelFn = void (*)(PassArgs);

template <bool INV, int... LRs>
constexpr auto outer_table(std::integer_sequence<int, LRs...>) {
    return std::array<KernelFn, sizeof...(LRs)>{&k_outer_pass<INV, LRs + 1>...};
}
template <int... LRs>
constexpr auto first_table(std::integer_sequence<int, LRs...>) {
    return std::array<KernelFn, sizeof...(LRs)>{&k_outer_first<LRs + 1>...};
}
template <int... LRs>
constexpr auto mid_table(std::integer_sequence<int, LRs...>) {
    return std::array<KernelFn, sizeof...(LRs)>{&k_mid_pass<LRs + 4, false>...};
}
template <int... LRs>
constexpr auto mid_limb_table(std::integer_sequence<int, LRs...>) {
    return std::array<KernelFn, sizeof...(LRs)>{&k_mid_pass<LRs + 4, true>...};
}
template <int... LRs>
constexpr auto pair_table(std::integer_sequence<int, LRs...>) {
    return std::array<KernelFn, sizeof...(LRs)>{&k_mid_pair<LRs + 4>...};
}

// Batched launches put one job per grid row; groups beyond the row limit run as several launches.
constexpr uint32_t kMaxGridY = 65535;

hipError_t launch_outer_pass(const PassArgs& a, bool inv, uint32_t njobs, uint32_t nops,
                             hipStream_t st) {
    static constexpr auto fwd = outer_table<false>(std::make_integer_sequence<int, 8>{});
    static constexpr auto bwd = outer_table<true>(std::make_integer_sequence<int, 8>{});
    static constexpr auto first = first_table(std::make_integer_sequence<int, 8>{});
    if (a.log_r < 1 || a.log_r > 8 || a.log_r + a.log_c != 12) return hipErrorInvalidValue;
    const uint32_t ntiles = 1u << (a.log_n - a.log_r - a.log_c);
    // a first forward pass (from limbs) handles the three primes inside each block
    const bool fused = !inv && (a.flags & kFromLimbs);
    (void)hipGetLastError();  // report this launch's error, not an earlier ignored one
    for (uint32_t j0 = 0; j0 < njobs; j0 += kMaxGridY) {  // one grid row per job
        PassArgs b = a;
        if (a.jobs) b.jobs = a.jobs + j0;
        if (j0) b.clear = nullptr;  // the first launch clears
        dim3 grid(ntiles, std::min(kMaxGridY, njobs - j0), fused ? nops : kNumPrimes * nops);
        hipLaunchKernelGGL((fused ? first : inv ? bwd : fwd)[a.log_r - 1], grid, dim3(kThreads), 0, st, b);
    }
    return hipGetLastError();
}

hipError_t launch_mid_pass(const PassArgs& a, uint32_t njobs, hipStream_t st) {
    static constexpr auto tab = mid_table(std::make_integer_sequence<int, 9>{});
    static constexpr auto ltab = mid_limb_table(std::make_integer_sequence<int, 9>{});
    static constexpr auto ptab = pair_table(std::make_integer_sequence<int, 9>{});
    if (a.log_r < 4 || a.log_r > 12) return hipErrorInvalidValue;
    const uint32_t nblk = 1u << (a.log_n - a.log_r);
    const uint32_t threads = (1u << a.log_r) / kE;
    (void)hipGetLastError();  // report this launch's error, not an earlier ignored one
    for (uint32_t j0 = 0; j0 < njobs; j0 += kMaxGridY) {
        PassArgs b = a;
        if (a.jobs) b.jobs = a.jobs + j0;
        if (j0) b.clear = nullptr;  // the first launch clears
        // plain products and squares: two tiles per block when the launch still has at least
        // 256 tile pairs per prime (small launches keep one tile per block for parallelism)
        const uint32_t nj = std::min(kMaxGridY, njobs - j0);
        const bool pair = (a.flags & ~kSquare) == 0 && nblk > 1 && (nblk / 2u) * nj >= 256u;
        dim3 grid(pair ? nblk / 2 : nblk, nj, kNumPrimes);
        const auto& t = pair ? ptab : ((a.flags & kFromLimbs) ? ltab : tab);
        hipLaunchKernelGGL(t[a.log_r - 4], grid, dim3(threads), 0, st, b);
    }
    return hipGetLastError();
}

hipError_t launch_diag_accumulate(const uint32_t* spec_a, const uint32_t* spec_b, uint32_t ca,
                                  uint32_t cb, uint32_t n, uint32_t* diag, hipStream_t st) {
    if (ca == 0 || cb == 0 || n < 256) return hipErrorInvalidValue;
    dim3 grid(n / 256u, ca + cb - 1u, kNumPrimes);
    (void)hipGetLastError();  // report this launch's error, not an earlier ignored one
    hipLaunchKernelGGL(k_diag_accumulate, grid, dim3(256), 0, st, spec_a, spec_b, ca, cb, n, diag);
    return hipGetLastError();
}

hipError_t launch_garner_carry(const GarnerArgs& a, uint32_t max_out_len, uint32_t njobs,
                               bool chunked, hipStream_t st) {
    const uint32_t ntiles = (max_out_len + kCarryTile - 1u) / kCarryTile;

    (void)hipGetLastError();  // report this launch's error, not an earlier ignored one
    for (uint32_t j0 = 0; j0 < njobs; j0 += kMaxGridY) {
        GarnerArgs b = a;
        if (a.jobs) b.jobs = a.jobs + j0;
        dim3 grid(ntiles, std::min(kMaxGridY, njobs - j0), 1);
        hipLaunchKernelGGL(chunked ? k_garner_carry<true> : k_garner_carry<false>, grid, dim3(kGThreads),
                           0, st, b);
        hipLaunchKernelGGL(k_carry_fixup, dim3(1, grid.y, 1), dim3(kFixThreads), 0, st, b);
    }
    return hipGetLastError();
}


}  // namespace bi
