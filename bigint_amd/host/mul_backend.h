// The product backend below mul_digits: products of stripped limb arrays above the host
// threshold.  The library's backend is the MI355X engine (mul_backend_hip.cpp).
#pragma once

#include <cstdint>
#include <span>
#include <vector>

#include "bigint/bigint_base.h"

namespace bigint::mul::backend {

struct Job {
    std::span<const uint32_t> a, b;  // nonzero top limbs; b == a (same data) for a square
    bool square;
    uint32_t* out;                   // a.size() + b.size() limbs
    size_t out_len;                  // written: product length without leading zeros
};

// Each job's product into its out buffer; throws on failure.
void multiply(std::vector<Job>& jobs, int plan_mode);

// The top min(keep, |a| + |b|) limbs of a * b into out (keep >= 1); returns the length of the
// whole product without leading zeros.  Throws on failure.
auto multiply_top(std::span<const uint32_t> a, std::span<const uint32_t> b, bool square, size_t keep,
                  uint32_t* out) -> size_t;

// Install a plan table / cost model file; false if unsupported or unusable.
bool load_planner(const char* path);

}  // namespace bigint::mul::backend
