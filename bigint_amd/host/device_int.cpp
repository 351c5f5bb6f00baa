// DeviceInt: numbers held in the product backend's memory between batches (include/bigint/device.h).
#include "bigint/device.h"

#include <algorithm>
#include <chrono>
#include <span>

#include "bigint/mul.h"

#include "mul_backend.h"

namespace bigint {
namespace {

// Runs f and adds its wall time to the engine time (bigint::mul::engine_time_ms).
template <class F>
void timed(F&& f) {
    const auto t0 = std::chrono::steady_clock::now();
    f();
    mul::add_engine_time_ns((uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                                std::chrono::steady_clock::now() - t0).count());
}

}  // namespace

auto DeviceInt::upload(const std::vector<const BigInt*>& xs) -> std::vector<DeviceInt> {
    std::vector<DeviceInt> res(xs.size());
    size_t words = 0;
    for (const BigInt* x : xs)
        if (!x->is_zero()) words += x->get_data().size();
    if (words == 0) return res;
    auto buf = mul::backend::dev_alloc(words);
    auto* base = static_cast<uint32_t*>(buf.get());
    std::vector<std::span<const uint32_t>> src;
    size_t off = 0;
    for (size_t i = 0; i < xs.size(); ++i) {
        if (xs[i]->is_zero()) continue;
        const Digits& d = xs[i]->get_data();
        src.emplace_back(d.data(), d.size());
        res[i].buf_ = buf;
        res[i].limbs_ = base + off;
        res[i].len_ = d.size();
        res[i].neg_ = xs[i]->sign() < 0;
        off += d.size();
    }
    timed([&] { mul::backend::dev_upload(base, src); });
    return res;
}

auto DeviceInt::download(const std::vector<const DeviceInt*>& xs) -> std::vector<BigInt> {
    // one copy per number (the numbers of a batch need not be adjacent in the backend)
    std::vector<Digits> digits(xs.size());
    for (size_t i = 0; i < xs.size(); ++i) digits[i].resize(xs[i]->len_);
    for (size_t i = 0; i < xs.size(); ++i) {
        if (!xs[i]->len_) continue;
        timed([&] { mul::backend::dev_download({std::span<uint32_t>(digits[i].data(), digits[i].size())}, xs[i]->limbs_); });
    }
    std::vector<BigInt> res;
    res.reserve(xs.size());
    for (size_t i = 0; i < xs.size(); ++i) res.push_back(BigInt::from_digits(std::move(digits[i]), xs[i]->neg_));
    return res;
}

auto DeviceInt::mul_batch(const std::vector<std::pair<const DeviceInt*, const DeviceInt*>>& pairs,
                          int plan_mode) -> std::vector<DeviceInt> {
    std::vector<DeviceInt> res(pairs.size());
    size_t words = 0;
    for (const auto& [a, b] : pairs)
        if (!a->is_zero() && !b->is_zero()) words += a->len_ + b->len_;
    if (words == 0) return res;
    auto buf = mul::backend::dev_alloc(words);
    auto* base = static_cast<uint32_t*>(buf.get());
    std::vector<mul::backend::DevMul> jobs;
    jobs.reserve(pairs.size());
    size_t off = 0;
    for (size_t i = 0; i < pairs.size(); ++i) {
        const DeviceInt& a = *pairs[i].first;
        const DeviceInt& b = *pairs[i].second;
        if (a.is_zero() || b.is_zero()) continue;
        const bool square = &a == &b || a.limbs_ == b.limbs_;
        jobs.push_back({a.limbs_, a.len_, b.limbs_, b.len_, square, base + off});
        res[i].buf_ = buf;
        res[i].limbs_ = base + off;
        res[i].len_ = a.len_ + b.len_;
        res[i].neg_ = a.neg_ != b.neg_;
        off += a.len_ + b.len_;
    }
    timed([&] { mul::backend::dev_mul_batch(jobs, plan_mode); });
    return res;
}

auto DeviceInt::add_batch(const std::vector<std::pair<const DeviceInt*, const DeviceInt*>>& pairs)
    -> std::vector<DeviceInt> {
    std::vector<DeviceInt> res(pairs.size());
    size_t words = 0;
    for (const auto& [a, b] : pairs)
        if (!a->is_zero() && !b->is_zero()) words += std::max(a->len_, b->len_) + 1;
    std::shared_ptr<void> buf = words ? mul::backend::dev_alloc(words) : nullptr;
    auto* base = static_cast<uint32_t*>(buf.get());
    std::vector<mul::backend::DevAdd> jobs;
    std::vector<size_t> at;
    size_t off = 0;
    for (size_t i = 0; i < pairs.size(); ++i) {
        const DeviceInt& a = *pairs[i].first;
        const DeviceInt& b = *pairs[i].second;
        if (a.is_zero() || b.is_zero()) {
            res[i] = a.is_zero() ? b : a;  // shares the operand's storage
            continue;
        }
        jobs.push_back({a.limbs_, a.len_, b.limbs_, b.len_, a.neg_ != b.neg_, base + off, false, 0});
        at.push_back(i);
        res[i].buf_ = buf;
        res[i].limbs_ = base + off;
        off += std::max(a.len_, b.len_) + 1;
    }
    timed([&] { mul::backend::dev_add_batch(jobs); });
    for (size_t k = 0; k < jobs.size(); ++k) {
        DeviceInt& r = res[at[k]];
        const DeviceInt& a = *pairs[at[k]].first;
        const DeviceInt& b = *pairs[at[k]].second;
        r.len_ = jobs[k].out_len;
        r.neg_ = r.len_ && (jobs[k].swapped ? b.neg_ : a.neg_);
        if (!r.len_) {
            r.buf_.reset();
            r.limbs_ = nullptr;
        }
    }
    return res;
}

}  // namespace bigint
