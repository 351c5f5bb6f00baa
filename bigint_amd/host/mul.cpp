// mul_digits: the host side of the multiply seam (reference include/bigint/mul.h:223-318).
#include "bigint/mul.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <thread>

#include "bigint/bigint.h"
#include "bigint_hip.h"
#include "mul_backend.h"

namespace bigint::mul {
namespace {

std::atomic<uint64_t> g_gpu_products{0};
std::atomic<uint64_t> g_host_products{0};
std::atomic<uint64_t> g_backend_ns{0};

// backend::multiply with its wall time added to g_backend_ns
void timed_multiply(std::vector<backend::Job>& jobs, int plan_mode) {
    const auto t0 = std::chrono::steady_clock::now();
    backend::multiply(jobs, plan_mode);
    g_backend_ns.fetch_add((uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                               std::chrono::steady_clock::now() - t0).count(),
                           std::memory_order_relaxed);
}

// 64-bit words of the value held in limbs (what the reference's digits_to_vec64 produces after
// dropping leading zero words, mul.h:245-268).
auto words64(std::span<const uint32_t> x) -> size_type {
    size_type n = x.size();
    while (n > 0 && x[n - 1] == 0) --n;
    if (n == 0) return 0;
    const size_type bits = (n - 1) * DIGIT_BITS + (size_type)std::bit_width(x[n - 1]);
    return (bits + 63) / 64;
}

// Schoolbook on 28-bit limbs by columns.  Every partial product is below 2^56, so up to 256 of
// them sum below 2^64: a column is summed in blocks of at most 256 products in 64-bit lanes (the
// loop vectorizes), and each block sum is split at bit 28 into the column's own part and the
// part carried into the next column.  The host route only takes products of <= 16,384 bits
// (at most 586 terms per column), so the carried part stays below 2^40.  No scratch array: the
// product is written as it goes.  Built twice, for AVX2 and generic x86-64, and chosen once at
// run time (an explicit choice rather than an ifunc, whose resolver runs before sanitizer
// runtimes start).
__attribute__((always_inline)) inline void schoolbook_cols_body(const uint32_t* a, size_type na,
                                                                const uint32_t* b, size_type nb,
                                                                uint32_t* r) {
    uint64_t carry = 0;  // carry into column k
    for (size_type k = 0; k + 1 < na + nb; ++k) {
        const size_type lo = k + 1 > nb ? k + 1 - nb : 0, hi = std::min(k, na - 1);
        uint64_t up = 0;
        for (size_type i = lo; i <= hi;) {
            const size_type e = std::min(hi + 1, i + 256);
            uint64_t sum = 0;
            for (; i < e; ++i) sum += (uint64_t)a[i] * b[k - i];
            up += sum >> DIGIT_BITS;
            carry += sum & DIGIT_MASK;
        }
        r[k] = (uint32_t)carry & DIGIT_MASK;
        carry = (carry >> DIGIT_BITS) + up;
    }
    r[na + nb - 1] = (uint32_t)carry;
}

__attribute__((target("avx2"))) void schoolbook_cols_avx2(const uint32_t* a, size_type na,
                                                          const uint32_t* b, size_type nb, uint32_t* r) {
    schoolbook_cols_body(a, na, b, nb, r);
}

void schoolbook_cols_generic(const uint32_t* a, size_type na, const uint32_t* b, size_type nb, uint32_t* r) {
    schoolbook_cols_body(a, na, b, nb, r);
}

bool cpu_has_avx2() {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx2");
}

void schoolbook_cols(const uint32_t* a, size_type na, const uint32_t* b, size_type nb, uint32_t* r) {
    static const bool avx2 = cpu_has_avx2();
    if (avx2) schoolbook_cols_avx2(a, na, b, nb, r);
    else schoolbook_cols_generic(a, na, b, nb, r);
}

auto schoolbook(std::span<const uint32_t> a, std::span<const uint32_t> b) -> Digits {
    Digits r(a.size() + b.size());
    schoolbook_cols(a.data(), a.size(), b.data(), b.size(), r.data());
    return r;
}

}  // namespace

auto gpu_product_count() noexcept -> uint64_t { return g_gpu_products.load(); }
auto host_product_count() noexcept -> uint64_t { return g_host_products.load(); }
auto engine_time_ms() noexcept -> double { return (double)g_backend_ns.load() * 1e-6; }
void add_engine_time_ns(uint64_t ns) noexcept { g_backend_ns.fetch_add(ns, std::memory_order_relaxed); }

auto load_planner(const char* path) -> bool { return backend::load_planner(path); }

// End-to-end crossover of single products between the host schoolbook and the engine (bits, as
// the reference counts them), measured on the MI355X box by bigint_amd/examples/crossover.cpp:
// the schoolbook takes 48 us at 46,464 total bits and 99 us at 65,664, the engine 54-59 us there.
constexpr size_type kHostCrossoverBits = 49152;
static std::atomic<size_type> g_host_max_bits{kHostCrossoverBits};

auto host_max_total_bits() noexcept -> size_type { return g_host_max_bits.load(std::memory_order_relaxed); }
void set_host_max_total_bits(size_type bits) noexcept {
    g_host_max_bits.store(std::max(bits, HOST_MAX_TOTAL_BITS), std::memory_order_relaxed);
}

// A product the reference would send to its NTT creates the thread pool at its default size
// (first call wins), as the reference's get_pool() does inside ntt::mul.
static void note_reference_route(size_type total_bits) {
    if (total_bits > REF_FFT_MAX_TOTAL_BITS && total_bits <= REF_NTT_MAX_TOTAL_BITS) (void)thread_pool_size();
}

auto mul_limbs(std::span<const uint32_t> a, std::span<const uint32_t> b, bool square) -> Digits {
    if (square) b = a;
    note_reference_route((words64(a) + words64(b)) * 64);
    if ((words64(a) + words64(b)) * 64 <= host_max_total_bits()) {
        g_host_products.fetch_add(1, std::memory_order_relaxed);
        Digits r = schoolbook(a, b);
        while (r.size() > 1 && r.back() == 0) r.pop_back();
        return r;
    }
    // strip leading zero limbs: the engine requires a nonzero top limb
    size_type la = a.size(), lb = b.size();
    while (la > 1 && a[la - 1] == 0) --la;
    while (lb > 1 && b[lb - 1] == 0) --lb;
    Digits r(la + lb);
    std::vector<backend::Job> jobs{backend::Job{a.first(la), b.first(lb), square, r.data(), 0}};
    timed_multiply(jobs, BI_HIP_PLAN_TABLE);
    g_gpu_products.fetch_add(1, std::memory_order_relaxed);
    r.resize(jobs[0].out_len);
    return r;
}

auto mul_digits(const Digits& a, const Digits& b, size_type output_precision,
                int64_t* p_result_point_pos) -> Digits {
    const bool same = &a == &b;
    size_type off_a =
        output_precision == 0 || a.size() <= output_precision ? 0 : a.size() - output_precision;
    size_type off_b =
        output_precision == 0 || b.size() <= output_precision ? 0 : b.size() - output_precision;
    size_type tz_a = 0, tz_b = 0;
    while (a[off_a + tz_a] == 0) ++tz_a;
    if (same) tz_b = tz_a;
    else
        while (b[off_b + tz_b] == 0) ++tz_b;
    off_a += tz_a;
    off_b += tz_b;

    const std::span<const uint32_t> sa(a.data() + off_a, a.size() - off_a);
    const std::span<const uint32_t> sb(b.data() + off_b, b.size() - off_b);
    if (output_precision) {
        // Only the top output_precision limbs survive (below): on the engine route, fetch just
        // the top output_precision + 1 limbs of the product (its length is la + lb or
        // la + lb - 1) instead of the whole product.
        size_type la = sa.size(), lb = sb.size();
        while (la > 1 && sa[la - 1] == 0) --la;
        while (lb > 1 && sb[lb - 1] == 0) --lb;
        const size_type keep = output_precision + 1;
        note_reference_route((words64(sa) + words64(sb)) * 64);
        if ((words64(sa) + words64(sb)) * 64 > host_max_total_bits() && la + lb > keep) {
            // staging for the engine's output: every limb is written, so it is not zero-filled
            const std::unique_ptr<uint32_t[]> top(new uint32_t[keep]);
            const auto t0 = std::chrono::steady_clock::now();
            const size_type n = backend::multiply_top(sa.first(la), sb.first(lb), same, keep, top.get());
            g_backend_ns.fetch_add((uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                                       std::chrono::steady_clock::now() - t0).count(),
                                   std::memory_order_relaxed);
            g_gpu_products.fetch_add(1, std::memory_order_relaxed);
            // top[i] is product limb (la + lb - keep + i); keep limbs [n - precision, n), then
            // drop their trailing zero limbs
            size_type off_r = n - output_precision;
            size_type i = off_r - (la + lb - keep);
            while (top[i] == 0) {
                ++i;
                ++off_r;
            }
            Digits res(top.get() + i, top.get() + (i + (n - off_r)));
            if (p_result_point_pos) *p_result_point_pos -= (int64_t)(off_a + off_b + off_r);
            return res;
        }
    }
    Digits res = mul_limbs(sa, sb, same);
    while (res.size() > 1 && res.back() == 0) res.pop_back();

    if (output_precision) {
        size_type off_r = res.size() <= output_precision ? 0 : res.size() - output_precision;
        while (res[off_r] == 0) ++off_r;
        if (off_r) res.erase(res.begin(), res.begin() + (std::ptrdiff_t)off_r);
        if (p_result_point_pos) *p_result_point_pos -= (int64_t)(off_a + off_b + off_r);
    } else {
        res.insert(res.begin(), tz_a + tz_b, 0u);
    }
    return res;
}

}  // namespace bigint::mul

namespace bigint::mul {

auto mul_digits_batch(const std::vector<std::pair<const Digits*, const Digits*>>& pairs,
                      int plan_mode) -> std::vector<Digits> {
    std::vector<Digits> res(pairs.size());
    struct Item {
        size_t idx;
        size_type za, zb;
        std::span<const uint32_t> a, b;
        bool same;
    };
    // Each thread of the init_thread_pool takes a contiguous range of pairs: host-sized products
    // (the reference's brute range) are multiplied at once, the rest are collected for one engine
    // batch.  Every result slot is written by one thread only.
    const size_t npairs = pairs.size();
    const size_t nthreads =
        npairs < 8192 ? 1 : std::max<size_t>(1, std::min<size_t>(thread_pool_size(), npairs / 4096));
    std::vector<std::vector<Item>> big_parts(nthreads);
    std::vector<size_t> host_counts(nthreads, 0);
    auto range = [&](size_t t, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            const Digits& a = *pairs[i].first;
            const Digits& b = *pairs[i].second;
            const bool same = &a == &b;
            size_type za = 0, zb = 0;
            while (a[za] == 0) ++za;
            if (same) zb = za;
            else
                while (b[zb] == 0) ++zb;
            std::span<const uint32_t> sa(a.data() + za, a.size() - za), sb(b.data() + zb, b.size() - zb);
            size_type la = sa.size(), lb = sb.size();
            while (la > 1 && sa[la - 1] == 0) --la;
            while (lb > 1 && sb[lb - 1] == 0) --lb;
            sa = sa.first(la);
            sb = sb.first(lb);
            if ((words64(sa) + words64(sb)) * 64 <= HOST_MAX_TOTAL_BITS) {
                Digits r = schoolbook(sa, same ? sa : sb);
                while (r.size() > 1 && r.back() == 0) r.pop_back();
                if (za + zb) r.insert(r.begin(), za + zb, 0u);
                res[i] = std::move(r);
                ++host_counts[t];
            } else {
                note_reference_route((words64(sa) + words64(sb)) * 64);
                big_parts[t].push_back(Item{i, za, zb, sa, sb, same});
            }
        }
    };
    if (nthreads == 1) {
        range(0, 0, npairs);
    } else {
        std::vector<std::thread> workers;
        const size_t per = (npairs + nthreads - 1) / nthreads;
        for (size_t t = 0; t < nthreads; ++t)
            workers.emplace_back(range, t, std::min(npairs, t * per), std::min(npairs, (t + 1) * per));
        for (auto& w : workers) w.join();
    }
    size_t nhost = 0, nbig = 0;
    for (size_t t = 0; t < nthreads; ++t) {
        nhost += host_counts[t];
        nbig += big_parts[t].size();
    }
    g_host_products.fetch_add(nhost, std::memory_order_relaxed);
    if (nbig == 0) return res;
    std::vector<Item> big;
    big.reserve(nbig);
    for (auto& part : big_parts) big.insert(big.end(), part.begin(), part.end());
    // Sizing the result vectors touches (zero-fills) every page of the products: spread it over
    // the pool's threads, products interleaved so that large ones land on different threads.
    std::vector<backend::Job> jobs(big.size());
    auto size_range = [&](size_t t, size_t step) {
        for (size_t k = t; k < big.size(); k += step) {
            Item& g = big[k];
            Digits& r = res[g.idx];
            r.assign(g.za + g.zb + g.a.size() + g.b.size(), 0u);
            jobs[k] = backend::Job{g.a, g.same ? g.a : g.b, g.same, r.data() + g.za + g.zb, 0};
        }
    };
    size_t out_limbs = 0;
    for (const Item& g : big) out_limbs += g.a.size() + g.b.size();
    const size_t nsize = out_limbs < (size_t(1) << 20)
                             ? 1
                             : std::max<size_t>(1, std::min<size_t>(thread_pool_size(), big.size()));
    if (nsize == 1) {
        size_range(0, 1);
    } else {
        std::vector<std::thread> workers;
        for (size_t t = 0; t < nsize; ++t) workers.emplace_back(size_range, t, nsize);
        for (auto& w : workers) w.join();
    }
    timed_multiply(jobs, plan_mode);
    g_gpu_products.fetch_add(big.size(), std::memory_order_relaxed);
    for (size_t k = 0; k < big.size(); ++k) res[big[k].idx].resize(big[k].za + big[k].zb + jobs[k].out_len);
    return res;
}

}  // namespace bigint::mul
