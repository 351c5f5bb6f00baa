// Crossover of single products between the host schoolbook and the MI355X engine, end to end:
// host limbs in, host limbs out, as mul_digits sees them (the engine side includes its host <->
// device copies and launches).  For operand sizes 2^12 .. 2^24 bits (balanced, random limbs) it
// prints one JSON line per size with the median time of each route, the host route only while
// it stays below 50 ms, and finally the smallest total size from which the engine is faster at
// every measured size above it.
//
//   crossover [reps]
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bigint/bigint.h"
#include "bigint/mul.h"
#include "bigint_hip.h"

using namespace bigint;
using clk = std::chrono::steady_clock;

static Digits random_limbs(size_t n, uint64_t& s) {
    Digits d(n);
    for (auto& x : d) {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        x = (uint32_t)((z ^ (z >> 31)) & DIGIT_MASK);
    }
    d.back() |= 1u << 27;
    return d;
}

template <class F>
static double median_us(F&& f, int reps) {
    f();  // warm-up
    std::vector<double> t;
    for (int r = 0; r < reps; ++r) {
        const auto t0 = clk::now();
        f();
        t.push_back(std::chrono::duration<double, std::micro>(clk::now() - t0).count());
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

int main(int argc, char** argv) {
    const int reps = argc > 1 ? std::atoi(argv[1]) : 15;
    if (bi_hip_init(0) != BI_HIP_OK) {
        std::fprintf(stderr, "no engine\n");
        return 1;
    }
    uint64_t seed = 0x9E3779B97F4A7C15ull;
    bool host_alive = true;
    std::vector<std::pair<size_t, bool>> engine_wins;  // (total bits, engine faster)
    for (int h = 24; h <= 48; ++h) {  // operand bits 2^(h/2)
        const size_t bits = (size_t)std::exp2(h / 2.0);
        const size_t n = (bits + DIGIT_BITS - 1) / DIGIT_BITS;
        const Digits a = random_limbs(n, seed), b = random_limbs(n, seed);
        const size_t total = 2 * ((n * DIGIT_BITS + 63) / 64) * 64;
        std::vector<uint32_t> out(2 * n);
        size_t len = 0;
        const double gpu = median_us([&] {
            if (bi_hip_mul(a.data(), n, b.data(), n, 0, out.data(), out.size(), &len) != BI_HIP_OK) std::abort();
        }, reps);
        double host = -1;
        if (host_alive) {
            mul::set_host_max_total_bits(~size_t(0) >> 1);
            host = median_us([&] { (void)mul::mul_digits(a, b); }, reps);
            mul::set_host_max_total_bits(mul::HOST_MAX_TOTAL_BITS);
            host_alive = host < 50000;
            engine_wins.push_back({total, gpu < host});
        }
        std::printf("{\"operand_bits\": %zu, \"total_bits\": %zu, \"host_us\": %.1f, \"engine_us\": %.1f}\n",
                    bits, total, host, gpu);
        std::fflush(stdout);
    }
    size_t cross = 0;  // smallest total size from which the engine wins at every measured size
    for (size_t i = engine_wins.size(); i-- > 0;) {
        if (!engine_wins[i].second) break;
        cross = engine_wins[i].first;
    }
    std::printf("{\"crossover_total_bits\": %zu}\n", cross);
    return 0;
}
