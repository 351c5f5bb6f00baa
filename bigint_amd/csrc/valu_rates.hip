// Issue-rate microbenchmark of the 32-bit integer VALU forms the NTT kernels are built from
// (v_add_u32, v_mul_lo_u32, v_mul_hi_u32, v_mad_u64_u32) on gfx950.  Measurement support only:
// it is a separate library (libbigint_rates.so, used by bench.py --valu-rates), not part of the
// engine's ABI.  Every thread runs eight independent dependency chains and the grid keeps eight
// waves on every SIMD, so the rate is the issue rate, not the latency.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

// OP 0: 8 adds per iteration; 1: 8 mul_lo; 2: 8 mul_hi + 8 xor; 3: 8 mad_u64 + 8 xor.
template <int OP>
__global__ __launch_bounds__(256) void k_rates(uint32_t* out, uint32_t seed, int iters) {
    uint32_t a[8], b[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        a[i] = seed * (threadIdx.x + 1u + i) | 1u;
        b[i] = seed + 77u * i + blockIdx.x;
    }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if constexpr (OP == 0) {
                a[i] = a[i] + b[i];
            } else if constexpr (OP == 1) {
                a[i] = a[i] * b[i];
            } else if constexpr (OP == 2) {
                a[i] = __umulhi(a[i], b[i]) ^ b[i];
            } else {
                const uint64_t t = (uint64_t)a[i] * b[i] + b[i];
                a[i] = (uint32_t)t ^ (uint32_t)(t >> 32);
            }
        }
    }
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) r ^= a[i];
    out[blockIdx.x * 256u + threadIdx.x] = r;  // keeps the chains live
}

}  // namespace

// Times one launch of `blocks` x 256 threads, `iters` iterations, after one untimed launch.
// Returns 0 or a hipError_t value; *ms is the launch's event-timed duration.
extern "C" int bi_rates_run(int op, int blocks, int iters, float* ms) {
    if (op < 0 || op > 3 || blocks <= 0 || iters <= 0 || !ms) return (int)hipErrorInvalidValue;
    uint32_t* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc((void**)&out, (size_t)blocks * 256u * 4u);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto k = op == 0 ? k_rates<0> : op == 1 ? k_rates<1> : op == 2 ? k_rates<2> : k_rates<3>;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, 0, out, 12345u, iters);
        e = hipEventRecord(e0, 0);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, 0, out, 12345u, iters);
        e = hipEventRecord(e1, 0);
    }
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    if (e == hipSuccess) e = hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out) (void)hipFree(out);
    return (int)e;
}

// Compute units and peak clock of device 0 (for the per-CU, per-clock rate).
extern "C" int bi_rates_device(int* cus, int* clock_khz) {
    if (!cus || !clock_khz) return (int)hipErrorInvalidValue;
    hipError_t e = hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, 0);
    if (e == hipSuccess) e = hipDeviceGetAttribute(clock_khz, hipDeviceAttributeClockRate, 0);
    return (int)e;
}
