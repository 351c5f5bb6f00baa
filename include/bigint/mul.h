// The multiply seam of the host library (reference include/bigint/mul.h:208-318).
//
// mul_digits keeps the reference's contract: both inputs nonzero, squaring detected by object
// identity (&a == &b), optional precision truncation of operands and result with the point
// position adjusted.  Below it, products of at most 16,384 total bits (counted in 64-bit words,
// reference mul.h:80, 210-212) use a host schoolbook; every larger product goes to the MI355X
// engine through the C ABI of include/bigint_hip.h.  There is no CPU fallback for those: if the
// engine is unavailable, mul_digits throws.
#pragma once

#include <cstdint>
#include <span>
#include <stdexcept>

#include "bigint/bigint_base.h"

namespace bigint::mul {

// The reference's brute-force limit (mul.h:80): products of at most this many total bits
// (counted in 64-bit words, as the reference counts them) are schoolbook products there.
constexpr size_type HOST_MAX_TOTAL_BITS = 2 * 128 * 64;

// Single products (mul_digits, operator*) of at most host_max_total_bits() total bits are
// multiplied by the host schoolbook, larger ones by the engine.  The default is the crossover
// measured end to end on the MI355X box (host schoolbook against the engine with its host copies);
// set_host_max_total_bits changes it (never below HOST_MAX_TOTAL_BITS).  Batches
// (mul_digits_batch) keep HOST_MAX_TOTAL_BITS: a batch shares its launch costs among its products.
auto host_max_total_bits() noexcept -> size_type;
void set_host_max_total_bits(size_type bits) noexcept;

// The reference's route boundaries (mul.h:153, 192, 208-220): it multiplies products of
// REF_FFT_MAX_TOTAL_BITS < (|A| + |B|) * 64 <= REF_NTT_MAX_TOTAL_BITS with its threaded NTT, whose
// first call creates the thread pool at its default size (get_pool(),
// src/mul/ntt_multithread.cpp:287-290), after which init_thread_pool has no effect.  Products in
// that band fix the pool size here the same way, whatever route computes them.
constexpr size_type REF_FFT_MAX_TOTAL_BITS = 10485760;
constexpr size_type REF_NTT_MAX_TOTAL_BITS = ((size_type(1) << 26) - 2) * 28;

// Raised when the GPU engine reports an error other than out-of-memory.
class engine_error : public std::runtime_error {
public:
    engine_error(int code, const char* what) : std::runtime_error(what), code_(code) {}
    [[nodiscard]] auto code() const noexcept -> int { return code_; }

private:
    int code_;
};

// out = a * b (28-bit limbs, canonical or with leading zeros); the core product below the
// precision/zero-stripping bookkeeping.  `square` means b is a.
auto mul_limbs(std::span<const uint32_t> a, std::span<const uint32_t> b, bool square) -> Digits;

// Reference semantics of mul::mul_digits (mul.h:223-318).  Inputs must be nonzero.
auto mul_digits(const Digits& a, const Digits& b, size_type output_precision = 0,
                int64_t* p_result_point_pos = nullptr) -> Digits;

// Number of products routed to the GPU engine / multiplied on the host since start-up.
auto gpu_product_count() noexcept -> uint64_t;
auto host_product_count() noexcept -> uint64_t;
// Wall time spent inside the product backend (engine calls, including staging copies).
auto engine_time_ms() noexcept -> double;
// adds to engine_time_ms (the DeviceInt batches report their backend time here)
void add_engine_time_ns(uint64_t ns) noexcept;

// Install a tuned plan table and cost model (the binary file bigint_amd/planner.py writes) for
// the engine's plan modes 1 (table) and 2 (learned); false if the backend has no planner or the
// file is unusable.  Without one the engine plans by its built-in heuristic.
auto load_planner(const char* path) -> bool;

}  // namespace bigint::mul

namespace bigint::mul {

// Products of many independent pairs (a_i * b_i), each exactly what mul_digits(a_i, b_i) returns
// with precision 0 (b_i == a_i by address selects squaring).  Pairs above the host threshold go
// to the engine as one grouped launch set; this is how a product-tree level runs on the GPU
// (the reference multiplies a level one `operator*` at a time, examples/pi.cpp:45-58).
auto mul_digits_batch(const std::vector<std::pair<const Digits*, const Digits*>>& pairs,
                      int plan_mode = 1) -> std::vector<Digits>;

}  // namespace bigint::mul
