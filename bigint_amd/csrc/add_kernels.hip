// Batched signed additions of device-resident limb arrays (28-bit limbs), for product-tree
// levels that stay in HBM: P = P_l Q_r + R_l P_r of the binary splitting (examples/pi.cpp:45-58
// merges with `operator*` and `operator+`, include/bigint/bigint.h; magnitude add/sub in
// src/bigint.cpp:267-548).
//
// A job adds |X| + |Y| or subtracts the smaller magnitude from the larger.  Carries (borrows)
// lie in {0, 1}, so a run of limbs acts on them as a map {0, 1} -> {0, 1} (two bits: the
// carry-out for carry-in 0 and for carry-in 1); maps compose associatively.  Three passes:
//   1. k_add_maps:  every 4096-limb tile composes its threads' maps (for a subtraction both
//      |X| - |Y| and |Y| - |X|);
//   2. k_add_scan:  one block per job scans its tiles' maps into each tile's carry-in; for a
//      subtraction the direction whose final borrow is 0 is the larger-minus-smaller one;
//   3. k_add_apply: each tile recomputes its limbs with the known carry-in, writes them and
//      records its highest nonzero limb (atomic max), so the result length is exact.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "add.h"
#include "ntt_device.h"

namespace bi {
namespace {

constexpr uint32_t kThreadsA = 256;
constexpr uint32_t kLimbsPerThread = 16;
constexpr uint32_t kTileA = kThreadsA * kLimbsPerThread;  // 4096 limbs
static_assert(kTileA == kAddTileLimbs, "tile size");
constexpr uint32_t kIdent = 0b10u;  // c -> c

__device__ __forceinline__ uint32_t amap_apply(uint32_t m, uint32_t c) { return (m >> c) & 1u; }
// apply `first`, then `second`
__device__ __forceinline__ uint32_t amap_compose(uint32_t second, uint32_t first) {
    return amap_apply(second, amap_apply(first, 0u)) | (amap_apply(second, amap_apply(first, 1u)) << 1);
}

// Limb i of x (zero past len).
__device__ __forceinline__ uint32_t limb_at(const uint32_t* x, size_t len, size_t i) {
    return i < len ? gmem(x)[i] : 0u;
}

// One limb step.  mode 0: a + b + c (carry out of 28 bits); mode 1: a - b - c (borrow out).
__device__ __forceinline__ uint32_t step(uint32_t a, uint32_t b, uint32_t c, bool sub, uint32_t& out) {
    if (!sub) {
        const uint32_t s = a + b + c;
        out = s & kDigitMask;
        return s >> kDigitBits;
    }
    const uint32_t t = b + c;
    out = (a - t) & kDigitMask;
    return a < t ? 1u : 0u;
}

// The thread's map over its 16 limbs for a + b (sub false) or a - b (sub true).
__device__ __forceinline__ uint32_t thread_map(const uint32_t (&a)[kLimbsPerThread],
                                               const uint32_t (&b)[kLimbsPerThread], bool sub) {
    uint32_t c0 = 0u, c1 = 1u, o;
#pragma unroll
    for (uint32_t q = 0; q < kLimbsPerThread; ++q) {
        c0 = step(a[q], b[q], c0, sub, o);
        c1 = step(a[q], b[q], c1, sub, o);
    }
    return c0 | (c1 << 1);
}

// Inclusive composition of the block's thread maps in thread order; returns the exclusive
// prefix of this thread and leaves the block total in *total.
__device__ __forceinline__ uint32_t block_scan(uint32_t m, uint32_t* s_wave, uint32_t* total) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t inc = m;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc = amap_compose(inc, o);
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    uint32_t prefix = kIdent;
    for (uint32_t w = 0; w < wave; ++w) prefix = amap_compose(s_wave[w], prefix);
    uint32_t all = kIdent;
    for (uint32_t w = 0; w < kThreadsA / 64; ++w) all = amap_compose(s_wave[w], all);
    *total = all;
    uint32_t excl = __shfl_up(inc, 1, 64);
    if (lane == 0) excl = kIdent;
    __syncthreads();  // s_wave may be reused by the caller
    return amap_compose(excl, prefix);
}

// LDS index of tile limb i: one pad word per 32 limbs, so the 16-limb runs that threads read
// (thread t at 16 t) fall in distinct banks across each half-wave.
__device__ __forceinline__ uint32_t tpad(uint32_t i) { return i + (i >> 5); }
constexpr uint32_t kTileLds = kTileA + kTileA / 32;

// Limbs [t0, t0 + 4096) of x (zero past len) into the thread's 16-limb run a[]: read from global
// memory lane-contiguous (limb k * 256 + t) and regrouped through LDS.  Leaves s free on return.
__device__ __forceinline__ void load_run(const uint32_t* x, size_t len, size_t t0, uint32_t* s,
                                         uint32_t (&a)[kLimbsPerThread]) {
    const uint32_t tid = threadIdx.x;
#pragma unroll
    for (uint32_t k = 0; k < kLimbsPerThread; ++k) {
        const uint32_t i = k * kThreadsA + tid;
        s[tpad(i)] = limb_at(x, len, t0 + i);
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < kLimbsPerThread; ++q) a[q] = s[tpad(tid * kLimbsPerThread + q)];
    __syncthreads();
}

__device__ __forceinline__ void load_tile(const AddJob& j, size_t t0, uint32_t* s, uint32_t (&a)[kLimbsPerThread],
                                          uint32_t (&b)[kLimbsPerThread]) {
    load_run(j.x, j.lx, t0, s, a);
    load_run(j.y, j.ly, t0, s, b);
}

// grid.x = tile, grid.y = job.  maps[tile_base + t] = map (add) or xy | yx << 2 (subtract).
__global__ __launch_bounds__(kThreadsA) void k_add_maps(const AddJob* __restrict__ jobs) {
    const AddJob j = load_desc(jobs + blockIdx.y);
    if (blockIdx.x >= j.ntiles) return;
    __shared__ uint32_t s_wave[kThreadsA / 64];
    __shared__ uint32_t s_limbs[kTileLds];
    uint32_t a[kLimbsPerThread], b[kLimbsPerThread];
    load_tile(j, (size_t)blockIdx.x * kTileA, s_limbs, a, b);
    uint32_t total;
    if (!j.subtract) {
        block_scan(thread_map(a, b, false), s_wave, &total);
        if (threadIdx.x == 0) gmem(j.maps)[blockIdx.x] = total;
    } else {
        uint32_t t_xy, t_yx;
        block_scan(thread_map(a, b, true), s_wave, &t_xy);
        block_scan(thread_map(b, a, true), s_wave, &t_yx);
        if (threadIdx.x == 0) gmem(j.maps)[blockIdx.x] = t_xy | (t_yx << 2);
    }
}

// One block per job: carry_in[t] for every tile (add: bit 0; subtract: bit 0 for x - y, bit 1
// for y - x), j.flags[0] = 1 when a subtraction's larger magnitude is |Y|, j.flags[1] cleared
// for k_add_apply's length.
__global__ __launch_bounds__(kThreadsA) void k_add_scan(const AddJob* __restrict__ jobs) {
    const AddJob j = load_desc(jobs + blockIdx.x);
    __shared__ uint32_t s_wave[kThreadsA / 64];
    uint32_t run0 = kIdent, run1 = kIdent;  // composition of the tiles before this chunk
    for (uint32_t base = 0; base < j.ntiles; base += kThreadsA) {
        const uint32_t t = base + threadIdx.x;
        const uint32_t v = t < j.ntiles ? gmem(j.maps)[t] : (j.subtract ? (kIdent | (kIdent << 2)) : kIdent);
        uint32_t tot0, tot1 = kIdent;
        const uint32_t e0 = amap_compose(block_scan(v & 3u, s_wave, &tot0), run0);
        uint32_t e1 = kIdent;
        if (j.subtract) e1 = amap_compose(block_scan((v >> 2) & 3u, s_wave, &tot1), run1);
        if (t < j.ntiles) gmem(j.carry_in)[t] = amap_apply(e0, 0u) | (amap_apply(e1, 0u) << 1);
        run0 = amap_compose(tot0, run0);
        run1 = amap_compose(tot1, run1);
    }
    if (threadIdx.x == 0) {
        // the final borrow of |X| - |Y| is 1 exactly when |Y| > |X|
        gmem(j.flags)[0] = j.subtract ? amap_apply(run0, 0u) : 0u;
        gmem(j.flags)[1] = 0u;  // highest nonzero limb + 1, raised by k_add_apply
    }
}

// grid.x = tile, grid.y = job: the result limbs [0, width) of the tile (an add's width includes
// the limb that takes the final carry), and the result length in j.flags[1].
__global__ __launch_bounds__(kThreadsA) void k_add_apply(const AddJob* __restrict__ jobs) {
    const AddJob j = load_desc(jobs + blockIdx.y);
    if (blockIdx.x >= j.ntiles) return;
    __shared__ uint32_t s_wave[kThreadsA / 64];
    __shared__ uint32_t s_top;
    __shared__ uint32_t s_limbs[kTileLds];
    if (threadIdx.x == 0) s_top = 0u;
    const bool swap = j.subtract && gmem(j.flags)[0] != 0u;
    const size_t t0 = (size_t)blockIdx.x * kTileA;
    uint32_t a[kLimbsPerThread], b[kLimbsPerThread];
    if (swap) load_tile(j, t0, s_limbs, b, a);
    else load_tile(j, t0, s_limbs, a, b);
    uint32_t total;
    const uint32_t excl = block_scan(thread_map(a, b, j.subtract), s_wave, &total);
    const uint32_t cin_tile = (gmem(j.carry_in)[blockIdx.x] >> (swap ? 1u : 0u)) & 1u;
    uint32_t c = amap_apply(excl, cin_tile);
    uint32_t top = 0u;  // 1 + index within the tile of the highest nonzero limb
#pragma unroll
    for (uint32_t q = 0; q < kLimbsPerThread; ++q) {
        uint32_t o;
        c = step(a[q], b[q], c, j.subtract, o);
        const uint32_t i = threadIdx.x * kLimbsPerThread + q;
        s_limbs[tpad(i)] = o;
        if (o && t0 + i < j.width) top = i + 1u;
    }
    __syncthreads();
    // the tile's limbs below width, written lane-contiguous
#pragma unroll
    for (uint32_t k = 0; k < kLimbsPerThread; ++k) {
        const uint32_t i = k * kThreadsA + threadIdx.x;
        if (t0 + i < j.width) gmem(j.out)[t0 + i] = s_limbs[tpad(i)];
    }
    if (top) atomicMax(&s_top, top);
    __syncthreads();
    if (threadIdx.x == 0 && s_top) {
        const uint32_t len = blockIdx.x * kTileA + s_top;
        __hip_atomic_fetch_max(gmem(j.flags) + 1, len, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace

hipError_t launch_add_batch(const AddJob* d_jobs, uint32_t njobs, uint32_t max_tiles, hipStream_t st) {
    if (njobs == 0) return hipSuccess;
    (void)hipGetLastError();
    k_add_maps<<<dim3(max_tiles, njobs), kThreadsA, 0, st>>>(d_jobs);
    k_add_scan<<<dim3(njobs), kThreadsA, 0, st>>>(d_jobs);
    k_add_apply<<<dim3(max_tiles, njobs), kThreadsA, 0, st>>>(d_jobs);
    return hipGetLastError();
}

}  // namespace bi
