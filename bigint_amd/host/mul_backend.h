// The product backend below mul_digits: products of stripped limb arrays above the host
// threshold.  The library's backend is the MI355X engine (mul_backend_hip.cpp).
#pragma once

#include <cstdint>
#include <memory>
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

// ---- numbers kept in the backend's memory between batches (DeviceInt) ----
// words uint32 of backend memory, freed with the last copy of the pointer.  Throws on failure.
auto dev_alloc(size_t words) -> std::shared_ptr<void>;
// the ranges of src, concatenated, into dst; dst's words, in order, into the ranges of out
void dev_upload(uint32_t* dst, const std::vector<std::span<const uint32_t>>& src);
void dev_download(const std::vector<std::span<uint32_t>>& out, const uint32_t* src);
struct DevMul {
    const uint32_t* a;
    size_t la;
    const uint32_t* b;
    size_t lb;
    bool square;
    uint32_t* out;  // la + lb limbs written (the top one may be zero)
};
void dev_mul_batch(const std::vector<DevMul>& jobs, int plan_mode);
struct DevAdd {
    const uint32_t* x;
    size_t lx;
    const uint32_t* y;
    size_t ly;
    bool subtract;    // |x| - |y| or |y| - |x|, whichever is not negative; else |x| + |y|
    uint32_t* out;    // max(lx, ly) + 1 limbs of room
    bool swapped;     // written: subtract and |y| > |x|
    size_t out_len;   // written: limbs without leading zeros, 0 for zero
};
void dev_add_batch(std::vector<DevAdd>& jobs);

}  // namespace bigint::mul::backend
