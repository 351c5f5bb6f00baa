// π by the Chudnovsky series with binary splitting, on the MI355X engine.
//
// The arithmetic after the splitting is the reference's examples/pi.cpp:91-126 operation for
// operation (BigFloat Newton iteration for 1/sqrt(10005), the reciprocal of 13591409*Q + P,
// precision-truncated products, decimal printing), so the digits are the reference's.  The
// splitting itself runs level by level: the reference merges terms with a streaming binary
// counter (examples/pi.cpp:13-80), one `operator*` at a time; here blocks of 64 consecutive terms
// are merged first in fixed-size limb arrays (no BigInt per leaf or per merge), then every
// level's products go to the engine together (BigInt::mul_batch), with the blocks and each
// level's additions spread over the init_thread_pool threads.  The merge (P, Q, R) o (P', Q', R')
// = (P Q' + R P', Q Q', R R') is associative, so any bracketing gives the same P(1, r) and
// Q(1, r).
//
// usage: pi <decimal digits> [plan mode 0|1|2] [output file] [planner file]
#include <malloc.h>

#include <bit>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <thread>
#include <vector>

#include "bigint/bigint.h"
#include "bigint/device.h"

using namespace bigint;
using clk = std::chrono::steady_clock;

#ifndef PI_DEC_PIECE_LIMBS
#define PI_DEC_PIECE_LIMBS 256
#endif
#ifndef PI_DEVICE_MIN_R_LIMBS
#define PI_DEVICE_MIN_R_LIMBS 0
#endif

struct Term {
    BigInt P, Q, R;
};

// Runs make(lo, hi) on the init_thread_pool threads over [0, count) split into contiguous ranges
// (at least min_per_thread items each) and concatenates the results in order.
template <class F>
static std::vector<Term> parallel_terms(size_type count, F make, size_type min_per_thread = 4096) {
    const size_type nt =
        std::max<size_type>(1, std::min<size_type>(thread_pool_size(), count / min_per_thread));
    std::vector<std::vector<Term>> parts(nt);
    std::vector<std::thread> threads;
    const size_type per = (count + nt - 1) / nt;
    for (size_type t = 0; t < nt; ++t)
        threads.emplace_back([&, t] { parts[t] = make(std::min(count, t * per), std::min(count, (t + 1) * per)); });
    for (auto& th : threads) th.join();
    std::vector<Term> out;
    out.reserve(count);
    for (auto& p : parts)
        for (auto& x : p) out.push_back(std::move(x));
    return out;
}

// ---- the lowest levels: blocks of kBlock terms merged in fixed-size limb arrays ----------------

// Signed integer of at most kCap 28-bit limbs (n == 0 is zero).
struct Small {
    static constexpr uint32_t kCap = 384;
    uint32_t n = 0;
    bool neg = false;
    uint32_t d[kCap];
};

static void small_set(Small& x, uint128_t v, bool neg) {
    x.n = 0;
    for (; v; v >>= DIGIT_BITS) x.d[x.n++] = (uint32_t)(v & DIGIT_MASK);
    x.neg = neg && x.n;
}

static void small_mul_u64(Small& x, uint64_t v) {
    uint128_t c = 0;
    for (uint32_t i = 0; i < x.n; ++i) {
        c += (uint128_t)x.d[i] * v;
        x.d[i] = (uint32_t)(c & DIGIT_MASK);
        c >>= DIGIT_BITS;
    }
    for (; c; c >>= DIGIT_BITS) x.d[x.n++] = (uint32_t)(c & DIGIT_MASK);
    if (!v) x.n = 0;
}

// r = |a| |b| (r distinct from a and b; the caller sets the sign).  Column sums in 64 bits,
// flushed every 256 products (each below 2^56); the column loop vectorizes.  Built twice, for
// AVX2 and generic x86-64, and chosen once at run time.
__attribute__((always_inline)) inline void small_mul_body(Small& r, const Small& a, const Small& b) {
    r.neg = false;
    if (!a.n || !b.n) {
        r.n = 0;
        return;
    }
    const uint32_t na = a.n, nb = b.n;
    if (na + nb > Small::kCap) throw std::length_error("pi: block term too large");
    uint64_t carry = 0;
    for (uint32_t k = 0; k + 1 < na + nb; ++k) {
        const uint32_t lo = k + 1 > nb ? k + 1 - nb : 0, hi = std::min(k, na - 1);
        uint64_t up = 0;
        for (uint32_t i = lo; i <= hi;) {
            const uint32_t e = std::min(hi + 1, i + 256);
            uint64_t sum = 0;
            for (; i < e; ++i) sum += (uint64_t)a.d[i] * b.d[k - i];
            up += sum >> DIGIT_BITS;
            carry += sum & DIGIT_MASK;
        }
        r.d[k] = (uint32_t)carry & DIGIT_MASK;
        carry = (carry >> DIGIT_BITS) + up;
    }
    r.d[na + nb - 1] = (uint32_t)carry;
    r.n = na + nb;
    while (r.n && r.d[r.n - 1] == 0) --r.n;
}

__attribute__((target("avx2"))) static void small_mul_avx2(Small& r, const Small& a, const Small& b) {
    small_mul_body(r, a, b);
}
static void small_mul_generic(Small& r, const Small& a, const Small& b) { small_mul_body(r, a, b); }

static void small_mul(Small& r, const Small& a, const Small& b) {
    static const bool avx2 = [] {
        __builtin_cpu_init();
        return __builtin_cpu_supports("avx2") != 0;
    }();
    if (avx2) small_mul_avx2(r, a, b);
    else small_mul_generic(r, a, b);
}

static int small_cmp_abs(const Small& a, const Small& b) {
    if (a.n != b.n) return a.n < b.n ? -1 : 1;
    for (uint32_t i = a.n; i-- > 0;)
        if (a.d[i] != b.d[i]) return a.d[i] < b.d[i] ? -1 : 1;
    return 0;
}

// r = x + y (signed; r distinct from x and y)
static void small_add(Small& r, const Small& x, const Small& y) {
    if (x.neg == y.neg) {
        const Small& a = x.n >= y.n ? x : y;
        const Small& b = x.n >= y.n ? y : x;
        if (a.n + 1 > Small::kCap) throw std::length_error("pi: block term too large");
        uint32_t c = 0;
        for (uint32_t i = 0; i < a.n; ++i) {
            const uint32_t t = a.d[i] + (i < b.n ? b.d[i] : 0u) + c;
            r.d[i] = t & DIGIT_MASK;
            c = t >> DIGIT_BITS;
        }
        r.n = a.n;
        if (c) r.d[r.n++] = c;
        r.neg = x.neg && r.n;
        return;
    }
    const bool x_big = small_cmp_abs(x, y) >= 0;
    const Small& a = x_big ? x : y;
    const Small& b = x_big ? y : x;
    uint32_t borrow = 0;
    for (uint32_t i = 0; i < a.n; ++i) {
        const uint32_t s = (i < b.n ? b.d[i] : 0u) + borrow;
        borrow = a.d[i] < s;
        r.d[i] = (a.d[i] + (borrow << DIGIT_BITS) - s) & DIGIT_MASK;
    }
    r.n = a.n;
    while (r.n && r.d[r.n - 1] == 0) --r.n;
    r.neg = a.neg && r.n;
}

struct SmallTerm {
    Small P, Q, R;
};

// Leaf i (examples/pi.cpp:26-44): Q = Q0 i^3, R = (2i-1)(6i-1)(6i-5), P = (-1)^i R (13591409 +
// 545140134 i), R and P with the same 128-bit intermediate arithmetic.
static void leaf_term(uint64_t i, SmallTerm& t) {
    constexpr uint64_t Q0 = 10939058860032000ull;  // 640320^3 / 24
    small_set(t.Q, Q0, false);
    small_mul_u64(t.Q, i);
    small_mul_u64(t.Q, i);
    small_mul_u64(t.Q, i);
    const auto rr = static_cast<uint128_t>(2 * i - 1) * (6 * i - 1) * (6 * i - 5);
    const uint64_t p0 = 13591409 + 545140134 * i;
    small_set(t.R, rr, false);
    small_set(t.P, rr * p0, i % 2 != 0);
}

// Terms 1..r merged in blocks of kBlock consecutive terms (the last block may be shorter): the
// first log2(kBlock) levels of the splitting without a BigInt per leaf or per merge.
constexpr uint64_t kBlock = 64;
constexpr int kBlockDepth = std::countr_zero(kBlock);

// (P, Q, R) of terms lo+1 .. hi: the merge (P, Q, R) o (P', Q', R') = (P Q' + R P', Q Q', R R').
// scratch holds two terms per level below this one (the recursion is at most kBlockDepth deep),
// so a block's merges allocate nothing.
static void block_term(uint64_t lo, uint64_t hi, SmallTerm& out, SmallTerm* scratch) {
    if (hi - lo == 1) {
        leaf_term(hi, out);
        return;
    }
    const uint64_t mid = lo + (hi - lo) / 2;
    SmallTerm& l = scratch[0];
    SmallTerm& h = scratch[1];
    block_term(lo, mid, l, scratch + 2);
    block_term(mid, hi, h, scratch + 2);
    Small t1, t2;
    small_mul(t1, l.P, h.Q);
    t1.neg = l.P.neg && t1.n;
    small_mul(t2, l.R, h.P);
    t2.neg = h.P.neg && t2.n;
    small_add(out.P, t1, t2);
    small_mul(out.Q, l.Q, h.Q);
    small_mul(out.R, l.R, h.R);
}

static BigInt to_bigint(const Small& x) {
    return BigInt::from_digits(Digits(x.d, x.d + x.n), x.neg);
}

static std::vector<Term> leaves(size_type r) {
    const size_type nblocks = (r + kBlock - 1) / kBlock;
    return parallel_terms(nblocks, [r](size_type lo, size_type hi) {
        std::vector<Term> t;
        t.reserve(hi - lo);
        SmallTerm st;
        const auto scratch = std::make_unique<SmallTerm[]>(2 * kBlockDepth);
        for (size_type b = lo; b < hi; ++b) {
            block_term(b * kBlock, std::min<uint64_t>(r, (b + 1) * kBlock), st, scratch.get());
            t.push_back(Term{to_bigint(st.P), to_bigint(st.Q), to_bigint(st.R)});
        }
        return t;
    }, 16);
}

// Level-wise binary splitting; returns (P, Q) of terms 1..r.  With PI_LEVELS set in the
// environment, one line per level goes to stderr (merges, largest Q in limbs, product and
// merge-addition milliseconds).
//
// After the block merge the tree stays in the product backend's memory (DeviceInt): each level
// is one product batch and one batch of signed additions, and only P(1, r) and Q(1, r) come
// back.  Host levels run before that only while R is below PI_DEVICE_MIN_R_LIMBS.
static std::pair<BigInt, BigInt> split(size_type r, int plan_mode, double& gpu_levels_ms) {
    const bool verbose = std::getenv("PI_LEVELS") != nullptr;
    const auto tl = clk::now();
    std::vector<Term> cur = leaves(r);
    if (verbose)
        std::cerr << "leaves: " << r << " terms in blocks of " << kBlock << ", ms "
                  << std::chrono::duration<double, std::milli>(clk::now() - tl).count() << "\n";
    // From here on the example allocates and frees vectors of 50-270 MB many times (the
    // downloads of P and Q, the Newton iterations, the final products, the decimal output).
    // glibc serves blocks above its mmap threshold with fresh mappings, so every such vector paid
    // for its pages again (about 0.6 s of page faults at 10^8 digits); keeping them in the heap
    // lets freed blocks be reused.  (Set after the leaves: their eight threads allocating many
    // small vectors run faster with the default policy.)
    mallopt(M_MMAP_THRESHOLD, 1 << 30);
    mallopt(M_TRIM_THRESHOLD, 1 << 30);
    int level = std::countr_zero(kBlock);  // the block merge did the levels below
    // The tree moves to the backend once every R has PI_DEVICE_MIN_R_LIMBS limbs: on the engine
    // at once (its batches beat host schoolbook on these levels); a backend without batching
    // (the CPU baseline build) waits until R R', a level's smallest product, is above the host
    // schoolbook's 16,384 bits (320 limbs per factor).
    auto all_big = [&] {
        for (const Term& t : cur)
            if (t.R.get_data().size() < PI_DEVICE_MIN_R_LIMBS) return false;
        return true;
    };
    while (cur.size() > 1 && !all_big()) {
        const size_type m = cur.size() / 2;
        const bool top = cur.size() <= 2;
        std::vector<std::pair<const BigInt*, const BigInt*>> jobs;
        jobs.reserve(4 * m);
        for (size_type k = 0; k < m; ++k) {
            const Term& l = cur[2 * k];
            const Term& h = cur[2 * k + 1];
            jobs.emplace_back(&l.P, &h.Q);
            jobs.emplace_back(&l.R, &h.P);
            jobs.emplace_back(&l.Q, &h.Q);
            if (!top) jobs.emplace_back(&l.R, &h.R);  // the final R is never used
        }
        const auto t0 = clk::now();
        std::vector<BigInt> prod = BigInt::mul_batch(jobs, plan_mode);
        const double pms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        gpu_levels_ms += pms;
        const auto t1 = clk::now();
        const size_type per = top ? 3 : 4;
        std::vector<Term> next = parallel_terms(m, [&](size_type lo, size_type hi) {
            std::vector<Term> t;
            t.reserve(hi - lo);
            for (size_type k = lo; k < hi; ++k) {
                BigInt P = std::move(prod[per * k]);
                P += prod[per * k + 1];
                t.push_back(Term{std::move(P), std::move(prod[per * k + 2]),
                                 top ? BigInt(0) : std::move(prod[per * k + 3])});
            }
            return t;
        });
        if (cur.size() % 2) next.push_back(std::move(cur.back()));
        if (verbose)
            std::cerr << "level " << level << ": merges " << m << ", largest Q limbs "
                      << next.front().Q.get_data().size() << ", products ms " << pms << ", merge ms "
                      << std::chrono::duration<double, std::milli>(clk::now() - t1).count() << "\n";
        ++level;
        cur = std::move(next);
    }
    if (cur.size() == 1) return {std::move(cur[0].P), std::move(cur[0].Q)};

    struct DTerm {
        DeviceInt P, Q, R;
    };
    const auto tu = clk::now();
    std::vector<DTerm> dc(cur.size());
    {
        std::vector<const BigInt*> xs;
        xs.reserve(3 * cur.size());
        for (const Term& t : cur) {
            xs.push_back(&t.P);
            xs.push_back(&t.Q);
            xs.push_back(&t.R);
        }
        std::vector<DeviceInt> d = DeviceInt::upload(xs);
        for (size_t i = 0; i < cur.size(); ++i) dc[i] = DTerm{std::move(d[3 * i]), std::move(d[3 * i + 1]), std::move(d[3 * i + 2])};
        cur.clear();
    }
    if (verbose)
        std::cerr << "upload of " << dc.size() << " terms, ms "
                  << std::chrono::duration<double, std::milli>(clk::now() - tu).count() << "\n";
    while (dc.size() > 1) {
        const size_type m = dc.size() / 2;
        const bool top = dc.size() <= 2;
        const size_type per = top ? 3 : 4;
        std::vector<std::pair<const DeviceInt*, const DeviceInt*>> jobs;
        jobs.reserve(per * m);
        for (size_type k = 0; k < m; ++k) {
            const DTerm& l = dc[2 * k];
            const DTerm& h = dc[2 * k + 1];
            jobs.emplace_back(&l.P, &h.Q);
            jobs.emplace_back(&l.R, &h.P);
            jobs.emplace_back(&l.Q, &h.Q);
            if (!top) jobs.emplace_back(&l.R, &h.R);
        }
        const auto t0 = clk::now();
        std::vector<DeviceInt> prod = DeviceInt::mul_batch(jobs, plan_mode);
        std::vector<std::pair<const DeviceInt*, const DeviceInt*>> adds(m);
        for (size_type k = 0; k < m; ++k) adds[k] = {&prod[per * k], &prod[per * k + 1]};
        std::vector<DeviceInt> sums = DeviceInt::add_batch(adds);
        const double pms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        gpu_levels_ms += pms;
        std::vector<DTerm> next(m);
        for (size_type k = 0; k < m; ++k)
            next[k] = DTerm{std::move(sums[k]), std::move(prod[per * k + 2]),
                            top ? DeviceInt() : std::move(prod[per * k + 3])};
        if (dc.size() % 2) next.push_back(std::move(dc.back()));
        if (verbose)
            std::cerr << "level " << level << " (device): merges " << m << ", largest Q limbs "
                      << next.front().Q.len() << ", products + additions ms " << pms << "\n";
        ++level;
        dc = std::move(next);
    }
    const auto td = clk::now();
    std::vector<BigInt> pq = DeviceInt::download({&dc[0].P, &dc[0].Q});
    if (verbose)
        std::cerr << "download of P, Q, ms "
                  << std::chrono::duration<double, std::milli>(clk::now() - td).count() << "\n";
    return {std::move(pq[0]), std::move(pq[1])};
}

int main(int argc, char** argv) {
    init_thread_pool(8);
    // Decimal output splits x = q 10^k + r down to pieces below this many limbs, which are
    // converted by repeated division (quadratic).  On the engine the split tree's products run
    // as batches, so pieces smaller than the library default (2000 limbs) are cheaper overall; a
    // build on a product backend without batching (the CPU baseline) keeps the default through
    // PI_DEC_PIECE_LIMBS.  The digits are the same for any threshold.
    BigInt::DEC_STRING_BRUTE_THRESHOLDS[1] = PI_DEC_PIECE_LIMBS;
    size_type dec_digits = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1000;
    const int plan_mode = argc > 2 ? std::atoi(argv[2]) : 1;
    const char* out_path = argc > 3 ? argv[3] : "pi.txt";
    const bool planner = argc > 4 && mul::load_planner(argv[4]);
    if (dec_digits < 2) dec_digits = 2;

    auto inner = static_cast<size_type>(
        std::ceil(static_cast<double>(dec_digits + 1) / DIGIT_BITS * std::log2(10)));
    inner = inner * 3 / 2;

    const auto t1 = clk::now();
    double levels_ms = 0;
    auto [P, Q] = split(dec_digits / 14 + 5, plan_mode, levels_ms);
    const auto t_split = clk::now();
    const double eng_split = mul::engine_time_ms();


    const bool verbose = std::getenv("PI_LEVELS") != nullptr;
    auto phase = [&, t = clk::now()](const char* name) mutable {
        if (!verbose) return;
        const auto now = clk::now();
        std::cerr << name << " ms " << std::chrono::duration<double, std::milli>(now - t).count() << "\n";
        t = now;
    };
    BigFloat pi(Q);
    pi *= 4270934400;
    BigFloat s(1.0 / std::sqrt(10005));
    {
        const BigFloat a(1.5);
        size_type cur = 52 / DIGIT_BITS;
        while (cur < inner) {
            cur *= 2;
            auto ss = BigFloat::mul(s, s, cur);
            ss *= 10005;
            ss >>= 1;
            s = BigFloat::mul(s, a - ss, cur);
        }
    }
    phase("1/sqrt(10005)");
    pi = BigFloat::mul(pi, s, inner);
    phase("Q * 4270934400 / sqrt(10005)");
    BigFloat den(13591409 * Q + P);
    phase("13591409 Q + P");
    BigFloat rden = den.reciprocal(inner);
    phase("reciprocal");
    pi = BigFloat::mul(pi, rden, inner);
    phase("final product");
    const auto t2 = clk::now();
    const double eng_compute = mul::engine_time_ms();

    {
        std::ofstream out(out_path, std::ios::out);
        out << print(pi, dec_digits, true) << '\n';
    }
    const auto t3 = clk::now();
    auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    std::cout << "{\"digits\": " << dec_digits << ", \"plan_mode\": " << plan_mode
              << ", \"planner_loaded\": " << (planner ? "true" : "false")
              << ", \"compute_ms\": " << ms(t1, t2) << ", \"split_ms\": " << ms(t1, t_split)
              << ", \"split_product_ms\": " << levels_ms << ", \"output_ms\": " << ms(t2, t3)
              << ", \"engine_ms\": {\"split\": " << eng_split << ", \"compute\": " << eng_compute
              << ", \"output\": " << mul::engine_time_ms() - eng_compute << "}"
              << ", \"gpu_products\": " << mul::gpu_product_count()
              << ", \"host_products\": " << mul::host_product_count() << "}" << std::endl;
    return 0;
}
