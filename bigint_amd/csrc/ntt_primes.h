// The three NTT primes and their generators, shared by the kernels and the host planner: the
// reference's own (include/bigint/ntt_base.h:14-16), P = {7*2^26 + 1, 15*2^27 + 1, 17*2^27 + 1}
// with generators {3, 31, 3}.
//
// P0 < 2^30 leaves room for unreduced butterfly values in [0, 4 P0); P1 < 2^31 for values in
// [0, 2 P1) where the next stage multiplies them; P2 > 2^31 keeps every value fully reduced, and
// its additions detect the carry out of 32 bits.
#pragma once

#include <stdint.h>

namespace bi {

constexpr int kNumPrimes = 3;
constexpr uint32_t kP[kNumPrimes] = {469762049u, 2013265921u, 2281701377u};
constexpr uint32_t kG[kNumPrimes] = {3u, 31u, 3u};
static_assert(kP[0] < 0x40000000u && kP[1] < 0x80000000u,
              "lazy butterflies need 4 P0 < 2^32 and 2 P1 < 2^32");

// True for a prime whose values may stay unreduced in [0, 2P) inside 32 bits.
constexpr bool lazy_prime(int i) { return kP[i] < 0x80000000u; }

}  // namespace bi
