// Device-side modular arithmetic for the three NTT primes (gfx950 / CDNA4).
//
// The 28-bit limb width and the primes are the reference's (include/bigint/ntt_base.h:14-20,
// ntt_primes.h); Montgomery reduction uses R = 2^32 as in include/bigint/ntt_base.h:40-71.  The
// prime index is a template parameter so every modulus and its -P^-1 mod 2^32 are immediates in
// the generated code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ntt_primes.h"

namespace bi {

// Global-memory views of device pointers.  Pointers read from job descriptors are generic to the
// compiler, which then emits flat loads and stores: those also count against the LDS wait
// counter, so every LDS wait would wait for outstanding global traffic too.  Accesses through
// these views compile to global_load / global_store.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

// Native vectors for 8- and 16-byte global accesses (HIP's uint2/uint4 classes do not copy across
// address spaces).
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Copy of a job descriptor read through a global view, one 32-bit word at a time; the index is
// uniform across the block, so the loads are scalar.
template <class T>
__device__ __forceinline__ T load_desc(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "descriptor size");
    T v;
    const auto* src = gmem(reinterpret_cast<const uint32_t*>(p));
    auto* dst = reinterpret_cast<uint32_t*>(&v);
#pragma unroll
    for (unsigned i = 0; i < sizeof(T) / 4; ++i) dst[i] = src[i];
    return v;
}

// -P^-1 mod 2^32.  Each P = k*2^m + 1 with 2m >= 32, so P*(P-2) = (P-1)^2 - 1 == -1 mod 2^32.
constexpr uint32_t kNegInvP[kNumPrimes] = {kP[0] - 2u, kP[1] - 2u, kP[2] - 2u};
constexpr uint32_t kDigitBits = 28;
constexpr uint32_t kDigitMask = (1u << kDigitBits) - 1u;
constexpr int kMaxLogN = 26;  // smallest 2-adic valuation of P-1 (P0: 26; P1, P2: 27)

static_assert(kP[0] * (kP[0] - 2u) == 0xFFFFFFFFu, "P0 Montgomery constant");
static_assert(kP[1] * (kP[1] - 2u) == 0xFFFFFFFFu, "P1 Montgomery constant");
static_assert(kP[2] * (kP[2] - 2u) == 0xFFFFFFFFu, "P2 Montgomery constant");

// a*b*2^-32 mod P for a, b < P.  t + u*P < P*(P + 2^32) < 2^64 for all three primes (P2 is the
// largest: 1.5e19 < 1.8e19), and the reduced value is below 1.54*P < 2^32, so one conditional
// subtraction finishes it; r - P wraps above r exactly when r < P, so an unsigned min selects.
template <int I>
__device__ __forceinline__ uint32_t mont_mul(uint32_t a, uint32_t b) {
    constexpr uint32_t p = kP[I];
    const uint64_t t = (uint64_t)a * b;
    const uint32_t u = (uint32_t)t * kNegInvP[I];
    const uint64_t s = t + (uint64_t)u * p;
    const uint32_t r = (uint32_t)(s >> 32);
    const uint32_t r2 = r - p;
    return r2 < r ? r2 : r;  // unsigned min: r - p wraps above r exactly when r < p
}

// t * 2^-32 mod P (Montgomery reduction of a 64-bit value) for t < 2^32 * P and
// t < 2^64 - 2^32 * P: then t + u*P does not overflow and (t + u*P) / 2^32 < 2P, so one
// conditional subtraction finishes it.  The caller guarantees the bound (P2: t < 8.6e18).
template <int I>
__device__ __forceinline__ uint32_t mont_redc(uint64_t t) {
    constexpr uint32_t p = kP[I];
    const uint32_t u = (uint32_t)t * kNegInvP[I];
    const uint64_t s = t + (uint64_t)u * p;
    const uint32_t r = (uint32_t)(s >> 32);
    const uint32_t r2 = r - p;
    return r2 < r ? r2 : r;
}

// Montgomery product without the final subtraction: for a < 4P, b < P and 4P < 2^32 (P0 only)
// the result lies in [0, 2P).
template <int I>
__device__ __forceinline__ uint32_t mont_mul_lazy(uint32_t a, uint32_t b) {
    static_assert(kP[I] < 0x40000000u, "lazy reduction needs 4P < 2^32");
    constexpr uint32_t p = kP[I];
    const uint64_t t = (uint64_t)a * b;
    const uint32_t u = (uint32_t)t * kNegInvP[I];
    return (uint32_t)((t + (uint64_t)u * p) >> 32);
}

// x mod 2P folded once: [0, 4P) -> [0, 2P) (4P < 2^32).
template <int I>
__device__ __forceinline__ uint32_t fold2p(uint32_t x) {
    constexpr uint32_t p2 = 2u * kP[I];
    const uint32_t y = x - p2;
    return y < x ? y : x;
}

// (a + b) mod P for a, b < P.
template <int I>
__device__ __forceinline__ uint32_t add_mod(uint32_t a, uint32_t b) {
    constexpr uint32_t p = kP[I];
    if constexpr (lazy_prime(I)) {
        const uint32_t s = a + b;  // < 2p < 2^32
        const uint32_t s2 = s - p;
        return s2 < s ? s2 : s;
    } else {
        // a + b may pass 2^32: subtract P when the sum carried out of 32 bits or is >= P
        uint32_t s, d;
        const bool carry = __builtin_add_overflow(a, b, &s);
        const bool borrow = __builtin_sub_overflow(s, p, &d);
        return (carry || !borrow) ? d : s;
    }
}

// (a - b) mod P for a, b < P.
template <int I>
__device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b) {
    constexpr uint32_t p = kP[I];
    if constexpr (lazy_prime(I)) {
        const uint32_t d = a - b;
        const uint32_t d2 = d + p;  // < 2p < 2^32 when no borrow
        return d2 < d ? d2 : d;
    } else {
        uint32_t d;
        const bool borrow = __builtin_sub_overflow(a, b, &d);
        return borrow ? d + p : d;  // d + p wraps back into [0, P) after a borrow
    }
}

}  // namespace bi
