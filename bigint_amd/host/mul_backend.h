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

// Install a plan table / cost model file; false if unsupported or unusable.
bool load_planner(const char* path);

}  // namespace bigint::mul::backend
