// π by the Chudnovsky series with binary splitting, on the MI355X engine.
//
// The arithmetic after the splitting is the reference's examples/pi.cpp:91-126 operation for
// operation (BigFloat Newton iteration for 1/sqrt(10005), the reciprocal of 13591409*Q + P,
// precision-truncated products, decimal printing), so the digits are the reference's.  The
// splitting itself runs level by level: the reference merges terms with a streaming binary
// counter (examples/pi.cpp:13-80), one `operator*` at a time; here every level's products go to
// the engine together (BigInt::mul_batch), and the leaves and each level's additions are spread
// over the init_thread_pool threads.  The merge (P, Q, R) o (P', Q', R') =
// (P Q' + R P', Q Q', R R') is associative, so any bracketing gives the same P(1, r) and Q(1, r).
//
// usage: pi <decimal digits> [plan mode 0|1|2] [output file] [planner file]
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <thread>
#include <vector>

#include "bigint/bigint.h"

using namespace bigint;
using clk = std::chrono::steady_clock;

struct Term {
    BigInt P, Q, R;
};

// Runs make(lo, hi) on the init_thread_pool threads over [0, count) split into contiguous ranges
// and concatenates the results in order.
template <class F>
static std::vector<Term> parallel_terms(size_type count, F make) {
    const size_type nt = std::max<size_type>(1, std::min<size_type>(thread_pool_size(), count / 4096));
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

// Leaf i (examples/pi.cpp:26-44): Q = Q0 i^3, R = (2i-1)(6i-1)(6i-5), P = (-1)^i R (13591409 +
// 545140134 i) with the same 128-bit intermediate arithmetic.
static std::vector<Term> leaves(size_type r) {
    constexpr uint64_t Q0 = 10939058860032000ull;  // 640320^3 / 24
    return parallel_terms(r, [](size_type lo, size_type hi) {
        std::vector<Term> t;
        t.reserve(hi - lo);
        for (size_type i = lo + 1; i <= hi; ++i) {
            BigInt q(Q0);
            q *= i;
            q *= i;
            q *= i;
            const auto rr = static_cast<uint128_t>(2 * i - 1) * (6 * i - 1) * (6 * i - 5);
            const uint64_t p0 = 13591409 + 545140134 * (uint64_t)i;
            BigInt p(rr * p0);
            if (i % 2) p.flip_sign();
            t.push_back(Term{std::move(p), std::move(q), BigInt(rr)});
        }
        return t;
    });
}

// Level-wise binary splitting; returns (P, Q) of terms 1..r.  With PI_LEVELS set in the
// environment, one line per level goes to stderr (merges, largest Q in limbs, product and
// merge-addition milliseconds).
static std::pair<BigInt, BigInt> split(size_type r, int plan_mode, double& gpu_levels_ms) {
    const bool verbose = std::getenv("PI_LEVELS") != nullptr;
    const auto tl = clk::now();
    std::vector<Term> cur = leaves(r);
    if (verbose)
        std::cerr << "leaves: " << r << " terms, ms "
                  << std::chrono::duration<double, std::milli>(clk::now() - tl).count() << "\n";
    int level = 0;
    while (cur.size() > 1) {
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
    return {std::move(cur[0].P), std::move(cur[0].Q)};
}

int main(int argc, char** argv) {
    init_thread_pool(8);
    // Decimal output splits x = q 10^k + r down to pieces below this many limbs, which are
    // converted by repeated division (quadratic).  The split tree's products run as batches, so
    // pieces smaller than the library default (2000 limbs) are cheaper overall; the digits are
    // the same for any threshold.
    BigInt::DEC_STRING_BRUTE_THRESHOLDS[1] = 256;
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
    pi = BigFloat::mul(pi, s, inner);
    pi = BigFloat::mul(pi, BigFloat(13591409 * Q + P).reciprocal(inner), inner);
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
