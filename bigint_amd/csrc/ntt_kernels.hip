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
//   * Garner CRT is fused with carry propagation, which runs as a decoupled look-back scan over
//     carry maps {0..3} -> {0..3} (ntt_multithread.cpp:384-442 does the same CRT and a serial
//     three-way overlapped carry).
//
// Twiddle convention: W[m + i] = w_{2m}^{bitrev_{log2 m}(i)} (Montgomery form), used by the
// butterfly of block i at the stage with m blocks.  Data stay in plain residue form: a Montgomery
// product with a Montgomery-form twiddle is an ordinary modular product.  The pointwise product
// is a Montgomery product, so the inverse output carries n * R^-1, removed in Garner.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ntt_device.h"
#include "ntt_plan.h"

namespace bi {

constexpr int kE = 16;                        // elements per thread in every pass
constexpr int kTile = 4096;                   // elements per tile per prime
constexpr int kThreads = kTile / kE;          // 256
constexpr int kLdsWords = kTile + kTile / 16; // data tile with one pad word per 16

__device__ __forceinline__ uint32_t lds_idx(uint32_t a) { return a + (a >> 4); }

// ---------------------------------------------------------------------------------------------
// Twiddle table for one tile: tw[k], k in [1, R), for the stage with tr = R / (2 * 2^floor(log2 k))
// rows of half-span:  tw[k] = W[k] * beta^tr  (beta = 1 for the first block of a pass).
template <int I>
__device__ void build_twiddles(uint32_t* __restrict__ tw, uint32_t* __restrict__ bpow,
                               const uint32_t* __restrict__ wtab, uint32_t beta, uint32_t log_r,
                               uint32_t nthreads) {
    const uint32_t R = 1u << log_r;
    const uint32_t tid = threadIdx.x;
    if (beta == 0) {  // beta == 1 marker: plain copy of the root table
        for (uint32_t k = tid; k < R; k += nthreads) tw[k] = wtab[k];
        return;
    }
    for (uint32_t e = tid; e < log_r; e += nthreads) {
        uint32_t b = beta;
        for (uint32_t j = 0; j < e; ++j) b = mont_mul<I>(b, b);
        bpow[e] = b;  // beta^(2^e)
    }
    __syncthreads();
    for (uint32_t k = tid; k < R; k += nthreads) {
        if (k == 0) continue;
        const uint32_t log_m = 31u - __builtin_clz(k);
        const uint32_t log_tr = log_r - 1u - log_m;
        tw[k] = mont_mul<I>(wtab[k], bpow[log_tr]);
    }
}

// ---------------------------------------------------------------------------------------------
// Register groups.  A group of EG = 2^e elements at rows base + q*tr_lo (q < EG) runs e radix-2
// stages.  Every thread holds 16 elements = 16/EG groups.  Row index of element (sg, q) is
// base[sg] + (q << log_trlo); the twiddle index of a butterfly whose lower row is r at the stage
// with log2(tr) = lt is (R >> (lt + 1)) + (r >> (lt + 1)).
template <int I, int EG>
__device__ __forceinline__ void fwd_group(uint32_t (&x)[kE], const uint32_t* __restrict__ tw,
                                          uint32_t log_r, uint32_t log_trlo,
                                          const uint32_t (&base)[kE / EG]) {
#pragma unroll
    for (int h = EG / 2; h >= 1; h >>= 1) {
        const uint32_t lt = log_trlo + (31u - __builtin_clz((uint32_t)h));
#pragma unroll
        for (int sg = 0; sg < kE / EG; ++sg) {
#pragma unroll
            for (int q = 0; q < EG; ++q) {
                if (q & h) continue;
                const uint32_t r = base[sg] + ((uint32_t)q << log_trlo);
                const uint32_t w = tw[(1u << (log_r - lt - 1u)) + (r >> (lt + 1u))];
                const uint32_t u = x[sg * EG + q];
                const uint32_t v = mont_mul<I>(x[sg * EG + q + h], w);
                x[sg * EG + q] = add_mod<I>(u, v);
                x[sg * EG + q + h] = sub_mod<I>(u, v);
            }
        }
    }
}

template <int I, int EG>
__device__ __forceinline__ void inv_group(uint32_t (&x)[kE], const uint32_t* __restrict__ tw,
                                          uint32_t log_r, uint32_t log_trlo,
                                          const uint32_t (&base)[kE / EG]) {
#pragma unroll
    for (int h = 1; h <= EG / 2; h <<= 1) {
        const uint32_t lt = log_trlo + (31u - __builtin_clz((uint32_t)h));
#pragma unroll
        for (int sg = 0; sg < kE / EG; ++sg) {
#pragma unroll
            for (int q = 0; q < EG; ++q) {
                if (q & h) continue;
                const uint32_t r = base[sg] + ((uint32_t)q << log_trlo);
                const uint32_t w = tw[(1u << (log_r - lt - 1u)) + (r >> (lt + 1u))];
                const uint32_t u = x[sg * EG + q];
                const uint32_t v = x[sg * EG + q + h];
                x[sg * EG + q] = add_mod<I>(u, v);
                x[sg * EG + q + h] = mont_mul<I>(sub_mod<I>(u, v), w);
            }
        }
    }
}

// Tile geometry: R = 2^log_r rows, C = 2^log_c columns, threads = R*C/16.
struct Geom {
    uint32_t log_r, log_c;
};

// Row base and column of sub-group sg for a group of EG elements with row stride tr_lo.
template <int EG>
__device__ __forceinline__ void group_layout(const Geom& g, uint32_t log_trlo, uint32_t nthreads,
                                             uint32_t (&base)[kE / EG], uint32_t (&col)[kE / EG]) {
    constexpr uint32_t le = (EG == 2) ? 1 : (EG == 4) ? 2 : (EG == 8) ? 3 : 4;
#pragma unroll
    for (int sg = 0; sg < kE / EG; ++sg) {
        const uint32_t s = threadIdx.x + nthreads * (uint32_t)sg;
        col[sg] = s & ((1u << g.log_c) - 1u);
        const uint32_t rest = s >> g.log_c;
        const uint32_t rl = rest & ((1u << log_trlo) - 1u);
        const uint32_t rh = rest >> log_trlo;
        base[sg] = (rh << (log_trlo + le)) + rl;
    }
}

// The group schedule of a tile: ceil(log_r / 4) groups from the top rows down; all groups have
// 4 stages except the last, which takes the remainder.
__device__ __forceinline__ int num_groups(uint32_t log_r) { return (int)((log_r + 3u) / 4u); }
__device__ __forceinline__ uint32_t group_stages(uint32_t log_r, int gi) {
    const uint32_t done = 4u * (uint32_t)gi;
    return (log_r - done) >= 4u ? 4u : (log_r - done);
}
__device__ __forceinline__ uint32_t group_trlo(uint32_t log_r, int gi) {
    return log_r - 4u * (uint32_t)gi - group_stages(log_r, gi);
}

// Element addressing callbacks: a tile element (row r, column c) <-> register slot.
// LDS exchange: write the current layout, read the next.
template <int EG>
__device__ __forceinline__ void lds_store(uint32_t* lds, const uint32_t (&x)[kE], const Geom& g,
                                          uint32_t log_trlo, uint32_t nthreads) {
    uint32_t base[kE / EG], col[kE / EG];
    group_layout<EG>(g, log_trlo, nthreads, base, col);
#pragma unroll
    for (int sg = 0; sg < kE / EG; ++sg)
#pragma unroll
        for (int q = 0; q < EG; ++q) {
            const uint32_t r = base[sg] + ((uint32_t)q << log_trlo);
            lds[lds_idx((r << g.log_c) + col[sg])] = x[sg * EG + q];
        }
}

template <int EG>
__device__ __forceinline__ void lds_load(const uint32_t* lds, uint32_t (&x)[kE], const Geom& g,
                                         uint32_t log_trlo, uint32_t nthreads) {
    uint32_t base[kE / EG], col[kE / EG];
    group_layout<EG>(g, log_trlo, nthreads, base, col);
#pragma unroll
    for (int sg = 0; sg < kE / EG; ++sg)
#pragma unroll
        for (int q = 0; q < EG; ++q) {
            const uint32_t r = base[sg] + ((uint32_t)q << log_trlo);
            x[sg * EG + q] = lds[lds_idx((r << g.log_c) + col[sg])];
        }
}

template <int I, bool INV>
__device__ __forceinline__ void run_group(int eg, uint32_t (&x)[kE], const uint32_t* tw,
                                          const Geom& g, uint32_t log_trlo, uint32_t nthreads) {
    switch (eg) {
#define BI_GROUP_CASE(EGV)                                                       \
    case EGV: {                                                                  \
        uint32_t base[kE / EGV], col[kE / EGV];                                  \
        group_layout<EGV>(g, log_trlo, nthreads, base, col);                     \
        if (INV) inv_group<I, EGV>(x, tw, g.log_r, log_trlo, base);              \
        else fwd_group<I, EGV>(x, tw, g.log_r, log_trlo, base);                  \
        break;                                                                   \
    }
        BI_GROUP_CASE(2)
        BI_GROUP_CASE(4)
        BI_GROUP_CASE(8)
        BI_GROUP_CASE(16)
#undef BI_GROUP_CASE
        default: break;
    }
}

__device__ __forceinline__ void exchange(uint32_t* lds, uint32_t (&x)[kE], const Geom& g,
                                         int eg_from, uint32_t trlo_from, int eg_to,
                                         uint32_t trlo_to, uint32_t nthreads) {
    __syncthreads();  // previous readers of lds are done
    switch (eg_from) {
        case 2: lds_store<2>(lds, x, g, trlo_from, nthreads); break;
        case 4: lds_store<4>(lds, x, g, trlo_from, nthreads); break;
        case 8: lds_store<8>(lds, x, g, trlo_from, nthreads); break;
        default: lds_store<16>(lds, x, g, trlo_from, nthreads); break;
    }
    __syncthreads();
    switch (eg_to) {
        case 2: lds_load<2>(lds, x, g, trlo_to, nthreads); break;
        case 4: lds_load<4>(lds, x, g, trlo_to, nthreads); break;
        case 8: lds_load<8>(lds, x, g, trlo_to, nthreads); break;
        default: lds_load<16>(lds, x, g, trlo_to, nthreads); break;
    }
}

// Full forward transform of the tile's R-point columns.  x arrives in the layout of group 0 and
// leaves in the layout of the last group.
template <int I>
__device__ void tile_forward(uint32_t (&x)[kE], uint32_t* lds, const uint32_t* tw, const Geom& g,
                             uint32_t nthreads) {
    const int ng = num_groups(g.log_r);
    for (int gi = 0; gi < ng; ++gi) {
        const uint32_t st = group_stages(g.log_r, gi);
        const uint32_t tl = group_trlo(g.log_r, gi);
        if (gi > 0) {
            const uint32_t pst = group_stages(g.log_r, gi - 1);
            exchange(lds, x, g, 1 << pst, group_trlo(g.log_r, gi - 1), 1 << st, tl, nthreads);
        }
        run_group<I, false>(1 << st, x, tw, g, tl, nthreads);
    }
}

// Inverse of tile_forward (times R): x arrives in the last group's layout and leaves in group 0's.
template <int I>
__device__ void tile_inverse(uint32_t (&x)[kE], uint32_t* lds, const uint32_t* tw, const Geom& g,
                             uint32_t nthreads) {
    const int ng = num_groups(g.log_r);
    for (int gi = ng - 1; gi >= 0; --gi) {
        const uint32_t st = group_stages(g.log_r, gi);
        const uint32_t tl = group_trlo(g.log_r, gi);
        if (gi < ng - 1) {
            const uint32_t pst = group_stages(g.log_r, gi + 1);
            exchange(lds, x, g, 1 << pst, group_trlo(g.log_r, gi + 1), 1 << st, tl, nthreads);
        }
        run_group<I, true>(1 << st, x, tw, g, tl, nthreads);
    }
}

// Global memory <-> registers in a given group layout.  Element (r, c) of tile (blk, cg) is at
// index blk * (R*S) + cg * C + c + r * S of a plane.  Limb input reads zero past `len`.
template <int EG, bool STORE>
__device__ __forceinline__ void global_io(uint32_t* plane, const uint32_t* limbs, uint32_t len,
                                          uint32_t (&x)[kE], const Geom& g, uint32_t log_trlo,
                                          uint32_t tile_base, uint32_t log_s, uint32_t nthreads) {
    uint32_t base[kE / EG], col[kE / EG];
    group_layout<EG>(g, log_trlo, nthreads, base, col);
#pragma unroll
    for (int sg = 0; sg < kE / EG; ++sg)
#pragma unroll
        for (int q = 0; q < EG; ++q) {
            const uint32_t r = base[sg] + ((uint32_t)q << log_trlo);
            const uint32_t idx = tile_base + col[sg] + (r << log_s);
            if (STORE) {
                plane[idx] = x[sg * EG + q];
            } else if (limbs) {
                x[sg * EG + q] = idx < len ? limbs[idx] : 0u;
            } else {
                x[sg * EG + q] = plane[idx];
            }
        }
}

template <bool STORE>
__device__ __forceinline__ void global_io_eg(int eg, uint32_t* plane, const uint32_t* limbs,
                                             uint32_t len, uint32_t (&x)[kE], const Geom& g,
                                             uint32_t log_trlo, uint32_t tile_base, uint32_t log_s,
                                             uint32_t nthreads) {
    switch (eg) {
        case 2: global_io<2, STORE>(plane, limbs, len, x, g, log_trlo, tile_base, log_s, nthreads); break;
        case 4: global_io<4, STORE>(plane, limbs, len, x, g, log_trlo, tile_base, log_s, nthreads); break;
        case 8: global_io<8, STORE>(plane, limbs, len, x, g, log_trlo, tile_base, log_s, nthreads); break;
        default: global_io<16, STORE>(plane, limbs, len, x, g, log_trlo, tile_base, log_s, nthreads); break;
    }
}

// ---------------------------------------------------------------------------------------------
// Outer (strided) pass.  grid.x = tiles, grid.y = job, grid.z = prime * nops + operand.
template <int I, bool INV>
__device__ void outer_pass_body(const PassArgs& a, const PassJob& job, uint32_t op) {
    __shared__ uint32_t lds[kLdsWords];
    __shared__ uint32_t tw[kTile];
    __shared__ uint32_t bpow[16];
    const Geom g{a.log_r, a.log_c};
    const uint32_t nthreads = kThreads;
    const uint32_t n = 1u << a.log_n;
    // tiles: ncg column groups per block, XCD-friendly order (consecutive tiles share lines)
    const uint32_t ntiles = gridDim.x;
    uint32_t t = blockIdx.x;
    if ((ntiles & 7u) == 0u) t = (t & 7u) * (ntiles >> 3) + (t >> 3);
    const uint32_t log_ncg = a.log_s - a.log_c;
    const uint32_t blk = t >> log_ncg;
    const uint32_t cg = t & ((1u << log_ncg) - 1u);
    const uint32_t tile_base = (blk << (a.log_r + a.log_s)) + (cg << a.log_c);

    const uint32_t* wt = a.wtab + ((INV ? kNumPrimes : 0) + I) * kTile;
    uint32_t beta = 0;
    if (a.beta) beta = a.beta[I * a.beta_stride + blk];
    build_twiddles<I>(tw, bpow, wt, blk == 0 ? 0u : beta, a.log_r, nthreads);

    uint32_t* plane = (op == 0 ? job.res_a : job.res_b) + (size_t)I * n;
    const uint32_t* limbs = nullptr;
    uint32_t len = 0;
    if (!INV && (a.flags & kFromLimbs)) {
        limbs = op == 0 ? job.limbs_a : job.limbs_b;
        len = op == 0 ? job.la : job.lb;
    }
    uint32_t x[kE];
    const int ng = num_groups(a.log_r);
    const int eg_first = 1 << group_stages(a.log_r, 0);
    const int eg_last = 1 << group_stages(a.log_r, ng - 1);
    const uint32_t tl_first = group_trlo(a.log_r, 0);
    if (!INV) {
        global_io_eg<false>(eg_first, plane, limbs, len, x, g, tl_first, tile_base, a.log_s, nthreads);
        __syncthreads();  // twiddles ready
        tile_forward<I>(x, lds, tw, g, nthreads);
        global_io_eg<true>(eg_last, plane, nullptr, 0, x, g, 0, tile_base, a.log_s, nthreads);
    } else {
        global_io_eg<false>(eg_last, plane, nullptr, 0, x, g, 0, tile_base, a.log_s, nthreads);
        __syncthreads();
        tile_inverse<I>(x, lds, tw, g, nthreads);
        global_io_eg<true>(eg_first, plane, nullptr, 0, x, g, tl_first, tile_base, a.log_s, nthreads);
    }
}

template <bool INV>
__global__ __launch_bounds__(kThreads) void k_outer_pass(PassArgs a) {
    const PassJob& job = a.jobs ? a.jobs[blockIdx.y] : a.job;
    const uint32_t nops = INV ? 1u : ((a.flags & kSquare) ? 1u : 2u);
    const uint32_t prime = blockIdx.z / nops, op = blockIdx.z % nops;
    switch (prime) {
        case 0: outer_pass_body<0, INV>(a, job, op); break;
        case 1: outer_pass_body<1, INV>(a, job, op); break;
        default: outer_pass_body<2, INV>(a, job, op); break;
    }
}

// ---------------------------------------------------------------------------------------------
// Block pass with the pointwise product: contiguous blocks of B = 2^log_r points (C = S = 1).
// Forward-transform the block of A (and of B unless squaring), multiply pointwise, inverse
// transform the block, store into A's plane.  grid.x = n / B, grid.y = job, grid.z = prime.
template <int I>
__device__ void mid_pass_body(const PassArgs& a, const PassJob& job) {
    __shared__ uint32_t lds[kLdsWords];
    __shared__ uint32_t twf[kTile];
    __shared__ uint32_t twi[kTile];
    __shared__ uint32_t bpow[2][16];
    const Geom g{a.log_r, 0};
    const uint32_t nthreads = blockDim.x;
    const uint32_t n = 1u << a.log_n;
    const uint32_t blk = blockIdx.x;
    const uint32_t tile_base = blk << a.log_r;

    const uint32_t* wf = a.wtab + I * kTile;
    const uint32_t* wi = a.wtab + (kNumPrimes + I) * kTile;
    uint32_t bf = 0, bi_ = 0;
    if (a.beta && blk != 0) {
        bf = a.beta[I * a.beta_stride + blk];
        bi_ = a.beta_inv[I * a.beta_stride + blk];
    }
    build_twiddles<I>(twf, bpow[0], wf, bf, a.log_r, nthreads);
    build_twiddles<I>(twi, bpow[1], wi, bi_, a.log_r, nthreads);

    const bool from_limbs = (a.flags & kFromLimbs) != 0;
    const bool square = (a.flags & kSquare) != 0;
    uint32_t* pa = job.res_a + (size_t)I * n;
    uint32_t* pb = job.res_b ? job.res_b + (size_t)I * n : nullptr;
    const int ng = num_groups(a.log_r);
    const int eg_first = 1 << group_stages(a.log_r, 0);
    const uint32_t tl_first = group_trlo(a.log_r, 0);

    uint32_t xa[kE], xb[kE];
    global_io_eg<false>(eg_first, pa, from_limbs ? job.limbs_a : nullptr, job.la, xa, g, tl_first,
                        tile_base, 0, nthreads);
    __syncthreads();  // twiddles ready
    tile_forward<I>(xa, lds, twf, g, nthreads);
    if (!square) {
        global_io_eg<false>(eg_first, pb, from_limbs ? job.limbs_b : nullptr, job.lb, xb, g,
                            tl_first, tile_base, 0, nthreads);
        tile_forward<I>(xb, lds, twf, g, nthreads);
#pragma unroll
        for (int k = 0; k < kE; ++k) xa[k] = mont_mul<I>(xa[k], xb[k]);
    } else {
#pragma unroll
        for (int k = 0; k < kE; ++k) xa[k] = mont_mul<I>(xa[k], xa[k]);
    }
    (void)ng;
    tile_inverse<I>(xa, lds, twi, g, nthreads);
    global_io_eg<true>(eg_first, pa, nullptr, 0, xa, g, tl_first, tile_base, 0, nthreads);
}

__global__ __launch_bounds__(kThreads) void k_mid_pass(PassArgs a) {
    const PassJob& job = a.jobs ? a.jobs[blockIdx.y] : a.job;
    switch (blockIdx.z) {
        case 0: mid_pass_body<0>(a, job); break;
        case 1: mid_pass_body<1>(a, job); break;
        default: mid_pass_body<2>(a, job); break;
    }
}

// ---------------------------------------------------------------------------------------------
// Garner CRT + carry propagation (decoupled look-back).
//
// Coefficient j of the convolution is recovered from the three inverse-transform outputs y_p
// (which carry a factor n * R^-1) as x_j < 2^84 and split into 28-bit pieces d0, d1, d2.
// Output limb i = d0_i + d1_{i-1} + d2_{i-2} + carry_i, the weights of
// ntt_multithread.cpp:413-442.  Each sum is below 3*2^28, so carry_i is in {0,1,2,3} and the
// carry-in -> carry-out behaviour of any run of limbs is a map {0..3} -> {0..3}, stored as four
// 2-bit fields.  Maps compose associatively, which makes the carry chain a scan.

__device__ __forceinline__ uint32_t map_apply(uint32_t m, uint32_t c) { return (m >> (2u * c)) & 3u; }
// (second after first): apply `first`, then `second`
__device__ __forceinline__ uint32_t map_compose(uint32_t second, uint32_t first) {
    uint32_t r = 0;
#pragma unroll
    for (uint32_t c = 0; c < 4; ++c) r |= map_apply(second, map_apply(first, c)) << (2u * c);
    return r;
}
constexpr uint32_t kMapIdentity = 0b11100100u;  // c -> c

struct Pieces {
    uint32_t d0, d1, d2;
};

__device__ __forceinline__ Pieces garner(const GarnerJob& jb, const GarnerConsts& k, uint32_t j) {
    if (j >= jb.n) return {0u, 0u, 0u};
    const uint32_t y0 = jb.y[j], y1 = jb.y[jb.n + j], y2 = jb.y[2u * jb.n + j];
    const uint32_t c0 = mont_mul<0>(y0, jb.kscale[0]);
    const uint32_t c1 = mont_mul<1>(y1, jb.kscale[1]);
    const uint32_t c2 = mont_mul<2>(y2, jb.kscale[2]);
    const uint32_t v0 = c0;
    const uint32_t v1 = mont_mul<1>(sub_mod<1>(c1, v0), k.inv01);
    const uint32_t t = mont_mul<2>(sub_mod<2>(c2, v0), k.inv02);
    const uint32_t v2 = mont_mul<2>(sub_mod<2>(t, v1), k.inv12);
    const uint64_t inner = (uint64_t)v1 + (uint64_t)kP[1] * v2;  // < 2^62.1
    const unsigned __int128 x = (unsigned __int128)inner * kP[0] + v0;
    const uint64_t lo = (uint64_t)x;
    const uint64_t hi = (uint64_t)(x >> 64);
    Pieces p;
    p.d0 = (uint32_t)lo & kDigitMask;
    p.d1 = (uint32_t)(lo >> 28) & kDigitMask;
    p.d2 = (uint32_t)((lo >> 56) | (hi << 8)) & kDigitMask;
    return p;
}

constexpr uint32_t kCarryLimbsPerThread = 16;
constexpr uint32_t kCarryTile = kThreads * kCarryLimbsPerThread;  // 4096 limbs per tile

__global__ __launch_bounds__(kThreads) void k_garner_carry(GarnerArgs a) {
    const GarnerJob& jb = a.jobs ? a.jobs[blockIdx.y] : a.job;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave[kThreads / 64];
    __shared__ uint32_t s_carry_in;
    if (threadIdx.x == 0) s_tile = atomicAdd(jb.counter, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t ntiles = (jb.out_len + kCarryTile - 1u) / kCarryTile;
    if (tile >= ntiles) return;

    const uint32_t i0 = tile * kCarryTile + threadIdx.x * kCarryLimbsPerThread;
    uint32_t s[kCarryLimbsPerThread];
    {
        Pieces pm2 = i0 >= 2 ? garner(jb, a.k, i0 - 2) : Pieces{0u, 0u, 0u};
        Pieces pm1 = i0 >= 1 ? garner(jb, a.k, i0 - 1) : Pieces{0u, 0u, 0u};
#pragma unroll
        for (uint32_t q = 0; q < kCarryLimbsPerThread; ++q) {
            const Pieces p = garner(jb, a.k, i0 + q);
            s[q] = p.d0 + pm1.d1 + pm2.d2;
            pm2 = pm1;
            pm1 = p;
        }
    }
    // this thread's carry map
    uint32_t m = kMapIdentity;
#pragma unroll
    for (uint32_t q = 0; q < kCarryLimbsPerThread; ++q) {
        uint32_t f = 0;
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) f |= ((s[q] + c) >> kDigitBits) << (2u * c);
        m = map_compose(f, m);
    }
    // inclusive scan within the wave
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t inc = m;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc = map_compose(inc, o);
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    uint32_t wave_prefix = kMapIdentity;  // composition of the waves before this one
    for (uint32_t w = 0; w < wave; ++w) wave_prefix = map_compose(s_wave[w], wave_prefix);
    uint32_t excl = __shfl_up(inc, 1, 64);
    if (lane == 0) excl = kMapIdentity;
    excl = map_compose(excl, wave_prefix);  // maps before this thread, within the tile

    if (threadIdx.x == 0) {
        uint32_t agg = kMapIdentity;
        for (uint32_t w = 0; w < kThreads / 64; ++w) agg = map_compose(s_wave[w], agg);
        uint32_t carry_in = 0;
        if (tile == 0) {
            __hip_atomic_store(&jb.status[0], (2u << 8) | map_apply(agg, 0u), __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&jb.status[tile], (1u << 8) | agg, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT);
            uint32_t acc = kMapIdentity;
            int32_t j = (int32_t)tile - 1;
            while (true) {
                uint32_t v;
                do {
                    v = __hip_atomic_load(&jb.status[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                } while ((v >> 8) == 0u);
                if ((v >> 8) == 2u) {
                    carry_in = map_apply(acc, v & 0xFFu);
                    break;
                }
                acc = map_compose(acc, v & 0xFFu);
                --j;
            }
            __hip_atomic_store(&jb.status[tile], (2u << 8) | map_apply(agg, carry_in),
                               __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_carry_in = carry_in;
    }
    __syncthreads();
    uint32_t carry = map_apply(excl, s_carry_in);
#pragma unroll
    for (uint32_t q = 0; q < kCarryLimbsPerThread; ++q) {
        const uint32_t v = s[q] + carry;
        const uint32_t i = i0 + q;
        if (i < jb.out_len) jb.out[i] = v & kDigitMask;
        carry = v >> kDigitBits;
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

}  // namespace bi

// ---------------------------------------------------------------------------------------------
// Launch wrappers (host side of this translation unit).
namespace bi {

hipError_t launch_outer_pass(const PassArgs& a, bool inv, uint32_t njobs, uint32_t nops,
                             hipStream_t st) {
    const uint32_t ntiles = 1u << (a.log_n - a.log_r - a.log_c);
    dim3 grid(ntiles, njobs, kNumPrimes * nops);
    if (inv) hipLaunchKernelGGL(k_outer_pass<true>, grid, dim3(kThreads), 0, st, a);
    else hipLaunchKernelGGL(k_outer_pass<false>, grid, dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_mid_pass(const PassArgs& a, uint32_t njobs, hipStream_t st) {
    const uint32_t nblk = 1u << (a.log_n - a.log_r);
    const uint32_t threads = (1u << a.log_r) / kE;
    dim3 grid(nblk, njobs, kNumPrimes);
    hipLaunchKernelGGL(k_mid_pass, grid, dim3(threads), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_garner_carry(const GarnerArgs& a, uint32_t max_out_len, uint32_t njobs,
                               hipStream_t st) {
    const uint32_t ntiles = (max_out_len + kCarryTile - 1u) / kCarryTile;
    dim3 grid(ntiles, njobs, 1);
    hipLaunchKernelGGL(k_garner_carry, grid, dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_fill(uint32_t* p, uint32_t v, uint32_t count, hipStream_t st) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((count + 255u) / 256u), dim3(256), 0, st, p, v, count);
    return hipGetLastError();
}

}  // namespace bi
