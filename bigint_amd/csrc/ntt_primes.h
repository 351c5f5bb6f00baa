// The three NTT primes and their generators, shared by the kernels and the host planner.
//
// The first two primes with their generators are the reference's (include/bigint/ntt_base.h:14-20).
// The third is 1811939329 = 27*2^26 + 1 (generator 13) in place of the reference's
// 2281701377 = 17*2^27 + 1: every prime is then below 2^31, which leaves room for unreduced
// butterfly outputs in [0, 2P) in all three transforms.  The product is the same integer: the
// prime product 2^90.47 still exceeds every convolution coefficient (< 2^90), so the CRT
// recovers the same convolution.
#pragma once

#include <stdint.h>

namespace bi {

constexpr int kNumPrimes = 3;
constexpr uint32_t kP[kNumPrimes] = {469762049u, 2013265921u, 1811939329u};
constexpr uint32_t kG[kNumPrimes] = {3u, 31u, 13u};
static_assert(kP[0] < 0x40000000u && kP[1] < 0x80000000u && kP[2] < 0x80000000u,
              "lazy butterflies and add_mod / sub_mod need 2P < 2^32 (4P for P0)");

}  // namespace bi
