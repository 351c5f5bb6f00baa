// Threaded AVX2 restatement of the reference's CPU NTT multiply — TEST/BENCH INFRASTRUCTURE ONLY.
//
// Used as the CPU baseline that bench.py times beside the GPU engine (cpu_baseline.kind =
// "port").  It follows the reference's algorithm and its runtime semantics so the timing is
// not a strawman:
//   - NTTThreadPool: a fixed pool, spin up to 5 ms then block (src/mul/ntt_multithread.cpp:184-203),
//     work split into blocks of max(1, total / (threads * 4)) claimed with fetch_add (37-42,
//     228-232), one barrier-synchronised phase per task (259-282); n in {0, 1} disables threads
//     (206, 216-218); the first init_thread_pool call wins (287-290, 407-409)
//   - serial thresholds: butterflies below 2 * 1024 points, mul/crt below 2048 / 1024 triples
//     (34, 85, 115, 144, 298, 317, 344, 390)
//   - 3-lane AVX2 Montgomery multiply and reduced add/sub on interleaved residues
//     (include/bigint/ntt_base.h:107-149)
//   - the pipeline of ntt::mul including both u64 <-> 28-bit repacks (ntt_multithread.cpp:444-530)
//     and mul_digits' repacks (include/bigint/mul.h:245-300)
// Built with the reference's Release flags -O3 -mavx -mavx2 -mfma -flto (CMakeLists.txt:34).
//   - the dispatcher _mul (mul.h:208-220) with its brute (mul.h:78-104) and single-threaded AVX2
//     FFT (src/mul/fft.cpp:281-386) routes below the NTT band; the SSA band above it is not
//     restated (no timed workload reaches it)
#include <immintrin.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <bit>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace {

using u128 = unsigned __int128;
constexpr uint32_t P[3] = {469762049u, 2013265921u, 2281701377u};
constexpr uint32_t G[3] = {3u, 31u, 3u};
constexpr uint32_t NIP[3] = {P[0] - 2u, P[1] - 2u, P[2] - 2u};  // -P^-1 mod 2^32
constexpr int DBITS = 28;
constexpr uint32_t DMASK = (1u << DBITS) - 1u;
constexpr int MAXLOG = 26;

uint32_t pw(uint32_t b, uint64_t e, uint32_t p) {
    uint64_t r = 1, x = b;
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (uint32_t)r;
}
uint32_t to_mont(int I, uint32_t x) { return (uint32_t)(((uint64_t)x << 32) % P[I]); }

struct Consts {
    std::array<uint32_t, 3> r_sq, one, inv01_etc;
    uint32_t steps[2][MAXLOG][3];
    uint32_t inv2[MAXLOG][3];
    Consts() {
        for (int I = 0; I < 3; ++I) {
            uint64_t rm = ((uint64_t)1 << 32) % P[I];
            r_sq[I] = (uint32_t)(rm * rm % P[I]);
            one[I] = (uint32_t)rm;
            uint32_t gi = pw(G[I], P[I] - 2, P[I]);
            for (int l = 1; l <= MAXLOG; ++l) {
                steps[0][l - 1][I] = to_mont(I, pw(G[I], (P[I] - 1) >> l, P[I]));
                steps[1][l - 1][I] = to_mont(I, pw(gi, (P[I] - 1) >> l, P[I]));
            }
            uint32_t h = pw(2, P[I] - 2, P[I]), acc = h;
            for (int l = 0; l < MAXLOG; ++l) {
                inv2[l][I] = to_mont(I, acc);
                acc = (uint32_t)((uint64_t)acc * h % P[I]);
            }
        }
    }
};
const Consts& K() {
    static Consts c;
    return c;
}

// ---- 3-lane AVX2 arithmetic on one interleaved residue triple -------------------------------
inline __m128i load3(const uint32_t* p) { return _mm_set_epi32(0, (int)p[2], (int)p[1], (int)p[0]); }
inline void store3(uint32_t* p, __m128i v) {
    _mm_storel_epi64((__m128i*)p, v);
    p[2] = (uint32_t)_mm_extract_epi32(v, 2);
}
inline __m128i vec3(const uint32_t* v) { return load3(v); }

inline __m128i mmul(__m128i a, __m128i b) {
    const __m256i pp = _mm256_set_epi64x(0, P[2], P[1], P[0]);
    const __m256i ps = _mm256_slli_epi64(pp, 32);
    const __m256i nip = _mm256_set_epi64x(0, NIP[2], NIP[1], NIP[0]);
    const __m256i flip = _mm256_set1_epi64x((long long)0x8000000000000000ull);
    __m256i t = _mm256_mul_epu32(_mm256_cvtepu32_epi64(a), _mm256_cvtepu32_epi64(b));
    __m256i u = _mm256_mul_epu32(_mm256_mul_epu32(t, nip), pp);
    t = _mm256_add_epi64(t, u);  // high halves hold the unreduced result
    // subtract P<<32 when t >= P<<32 (unsigned compare through the sign flip)
    __m256i lt = _mm256_cmpgt_epi64(_mm256_xor_si256(ps, flip), _mm256_xor_si256(t, flip));
    t = _mm256_sub_epi64(t, _mm256_andnot_si256(lt, ps));
    __m256i hi = _mm256_permutevar8x32_epi32(t, _mm256_set_epi32(7, 6, 5, 4, 7, 5, 3, 1));
    return _mm256_castsi256_si128(hi);
}
inline __m128i madd(__m128i a, __m128i b) {
    const __m128i p = _mm_set_epi32(0, (int)P[2], (int)P[1], (int)P[0]);
    const __m128i flip = _mm_set1_epi32((int)0x80000000u);
    __m128i room = _mm_sub_epi32(p, b);
    __m128i keep = _mm_cmplt_epi32(_mm_xor_si128(a, flip), _mm_xor_si128(room, flip));
    return _mm_sub_epi32(_mm_add_epi32(a, b), _mm_andnot_si128(keep, p));
}
inline __m128i msub(__m128i a, __m128i b) {
    const __m128i p = _mm_set_epi32(0, (int)P[2], (int)P[1], (int)P[0]);
    const __m128i flip = _mm_set1_epi32((int)0x80000000u);
    __m128i borrow = _mm_cmplt_epi32(_mm_xor_si128(a, flip), _mm_xor_si128(b, flip));
    return _mm_add_epi32(_mm_sub_epi32(a, b), _mm_and_si128(borrow, p));
}
inline __m128i mpow(__m128i b, uint64_t e) {
    __m128i r = vec3(K().one.data());
    while (e) {
        if (e & 1) r = mmul(r, b);
        b = mmul(b, b);
        e >>= 1;
    }
    return r;
}

// ---- pool -------------------------------------------------------------------------------------
class Pool {
public:
    explicit Pool(uint32_t n) : nthreads_(n), running_(n), valid_(n != 0 && n != 1) {
        if (!valid_) return;
        for (uint32_t i = 0; i < nthreads_; ++i) workers_.emplace_back([this] { loop(); });
    }
    ~Pool() {
        if (!valid_) return;
        stop_.store(true, std::memory_order_release);
        work_.store(true, std::memory_order_release);
        work_.notify_all();
        for (auto& t : workers_) t.join();
    }
    bool valid() const { return valid_; }
    uint32_t threads() const { return nthreads_; }

    // Run body(first, last) over [0, total) split in blocks of max(1, total / (threads * 4)).
    void parallel(size_t total, const std::function<void(size_t, size_t)>& body) {
        std::lock_guard<std::mutex> lk(mtx_);
        body_ = &body;
        total_ = total;
        block_ = std::max<size_t>(1, total / ((size_t)nthreads_ * 4));
        next_.store(0, std::memory_order_relaxed);
        work_.store(true, std::memory_order_release);
        work_.notify_all();
        while (running_.load(std::memory_order_acquire) != 0) _mm_pause();
        work_.store(false, std::memory_order_relaxed);
        done_.store(true, std::memory_order_release);
        done_.notify_all();
        while (running_.load(std::memory_order_acquire) != nthreads_) _mm_pause();
        done_.store(false, std::memory_order_relaxed);
    }

private:
    static void spin_then_block(std::atomic<bool>& flag) {
        const auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(5);
        do {
            std::this_thread::yield();
            for (int i = 0; i < 32; ++i) {
                if (flag.load(std::memory_order_acquire)) return;
                _mm_pause();
            }
        } while (std::chrono::steady_clock::now() < deadline);
        while (!flag.load(std::memory_order_acquire)) flag.wait(false, std::memory_order_acquire);
    }
    void loop() {
        for (;;) {
            spin_then_block(work_);
            if (stop_.load(std::memory_order_acquire)) return;
            for (;;) {
                const size_t b = next_.fetch_add(1, std::memory_order_relaxed);
                const size_t s = b * block_;
                if (s >= total_) break;
                (*body_)(s, std::min(s + block_, total_));
            }
            running_.fetch_sub(1, std::memory_order_release);
            spin_then_block(done_);
            running_.fetch_add(1, std::memory_order_release);
        }
    }
    uint32_t nthreads_;
    std::atomic<uint32_t> running_;
    bool valid_;
    std::atomic<bool> work_{false}, done_{false}, stop_{false};
    std::atomic<size_t> next_{0};
    size_t total_ = 0, block_ = 1;
    const std::function<void(size_t, size_t)>* body_ = nullptr;
    std::mutex mtx_;
    std::vector<std::thread> workers_;
};

Pool& pool(uint32_t n = 0) {
    static Pool p(n != 0 ? n : std::thread::hardware_concurrency());
    return p;
}

// ---- NTT stages --------------------------------------------------------------------------------
void mul_num(uint32_t* v, size_t pts, const std::array<uint32_t, 3>& k) {
    auto body = [&](size_t s, size_t e) {
        const __m128i kk = vec3(k.data());
        for (size_t i = s; i < e; ++i) store3(v + 3 * i, mmul(load3(v + 3 * i), kk));
    };
    if (pts < 2048 || !pool().valid()) body(0, pts);
    else pool().parallel(pts, body);
}

void mul_vec(uint32_t* v, const uint32_t* w, size_t pts) {
    auto body = [&](size_t s, size_t e) {
        for (size_t i = s; i < e; ++i) store3(v + 3 * i, mmul(load3(v + 3 * i), load3(w + 3 * i)));
    };
    if (pts < 2048 || !pool().valid()) body(0, pts);
    else pool().parallel(pts, body);
}

void bit_reverse(uint32_t* v, size_t n) {
    for (size_t i = 1, j = n >> 1; i < n - 1; ++i) {
        if (i < j)
            for (int u = 0; u < 3; ++u) std::swap(v[3 * i + u], v[3 * j + u]);
        size_t k = n >> 1;
        while (j >= k) {
            j -= k;
            k >>= 1;
        }
        j += k;
    }
}

void transform(uint32_t* v, size_t n, int rev) {
    if (n == 1) return;
    bit_reverse(v, n);
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t half = len / 2;
        const __m128i step = vec3(K().steps[rev][std::countr_zero(half)]);
        auto body = [&](size_t s, size_t e) {
            __m128i cur = vec3(K().one.data());
            if (size_t j0 = s & (half - 1); j0) cur = mpow(step, j0);
            for (size_t c = s; c < e; ++c) {
                const size_t j = c & (half - 1), i = (c - j) << 1;
                if (j == 0) cur = vec3(K().one.data());
                uint32_t* pa = v + 3 * (i + j);
                uint32_t* pb = v + 3 * (i + j + half);
                const __m128i a = load3(pa), b = mmul(cur, load3(pb));
                store3(pa, madd(a, b));
                store3(pb, msub(a, b));
                cur = mmul(cur, step);
            }
        };
        if (n < 2048 || !pool().valid()) body(0, n / 2);
        else pool().parallel(n / 2, body);
    }
    if (rev) {
        std::array<uint32_t, 3> s;
        std::memcpy(s.data(), K().inv2[std::countr_zero(n) - 1], 12);
        mul_num(v, n, s);
    }
}

void crt(uint32_t* v, size_t pts) {
    const uint32_t i01 = pw(P[0] % P[1], P[1] - 2, P[1]);
    const uint32_t i02 = pw(P[0] % P[2], P[2] - 2, P[2]);
    const uint32_t i12 = pw(P[1] % P[2], P[2] - 2, P[2]);
    auto body = [&](size_t s, size_t e) {
        for (size_t i = s; i < e; ++i) {
            uint32_t* q = v + 3 * i;
            const uint32_t v0 = q[0];
            const uint32_t d1 = q[1] >= v0 ? q[1] - v0 : q[1] + P[1] - v0;
            const uint32_t v1 = (uint32_t)((uint64_t)d1 * i01 % P[1]);
            const uint32_t d2 = q[2] >= v0 ? q[2] - v0 : q[2] + P[2] - v0;
            uint32_t v2 = (uint32_t)((uint64_t)d2 * i02 % P[2]);
            v2 = v2 >= v1 ? v2 - v1 : v2 + P[2] - v1;
            v2 = (uint32_t)((uint64_t)v2 * i12 % P[2]);
            u128 x = (u128)((uint64_t)P[0] * P[1]) * v2 + (uint64_t)P[0] * v1 + v0;
            q[0] = (uint32_t)x & DMASK;
            q[1] = (uint32_t)(x >> 28) & DMASK;
            q[2] = (uint32_t)(x >> 56) & DMASK;
        }
    };
    if (pts < 1024 || !pool().valid()) body(0, pts);
    else pool().parallel(pts, body);
}

// overlapped three-way sum + serial carry; returns trimmed length
size_t trim(std::vector<uint32_t>& v, size_t n) {
    if (n != 1) {
        uint32_t d = v[1] + v[3];
        v[1] = d & DMASK;
        uint32_t c = d >> DBITS;
        for (size_t i = 2, j = 6; i < n; ++i, j += 3) {
            d = v[j - 4] + v[j - 2] + v[j] + c;
            v[i] = d & DMASK;
            c = d >> DBITS;
        }
        d = v[3 * n - 4] + v[3 * n - 2] + c;
        v[n] = d & DMASK;
        c = d >> DBITS;
        d = v[3 * n - 1] + c;
        v[n + 1] = d & DMASK;
        v[n + 2] = d >> DBITS;
    }
    size_t i = n == 1 ? 2 : n + 2;
    while (v[i] == 0) --i;
    return i + 1;
}

void words_to_digits3(const uint64_t* src, size_t nw, std::vector<uint32_t>& dst) {
    u128 t = 0;
    int bits = 0;
    size_t i = 0;
    while (i < nw) {
        if (bits < DBITS) {
            t |= (u128)src[i++] << bits;
            bits += 64;
        }
        while (bits >= DBITS) {
            const uint32_t d = (uint32_t)t & DMASK;
            dst.push_back(d);
            dst.push_back(d);
            dst.push_back(d);
            t >>= DBITS;
            bits -= DBITS;
        }
    }
    while (t) {
        const uint32_t d = (uint32_t)t & DMASK;
        dst.push_back(d);
        dst.push_back(d);
        dst.push_back(d);
        t >>= DBITS;
    }
}

std::vector<uint64_t> ntt_mul(const std::vector<uint64_t>& A, const std::vector<uint64_t>& B,
                              bool same) {
    const size_t need = (A.size() * 64 + DBITS - 1) / DBITS + (B.size() * 64 + DBITS - 1) / DBITS - 1;
    const size_t n = std::bit_ceil(need);
    std::vector<uint32_t> va;
    va.reserve(3 * n + 16);
    words_to_digits3(A.data(), A.size(), va);
    va.resize(3 * n + 16);
    mul_num(va.data(), n, K().r_sq);
    transform(va.data(), n, 0);
    if (same) {
        mul_vec(va.data(), va.data(), n);
    } else {
        std::vector<uint32_t> vb;
        vb.reserve(3 * n + 16);
        words_to_digits3(B.data(), B.size(), vb);
        vb.resize(3 * n + 16);
        mul_num(vb.data(), n, K().r_sq);
        transform(vb.data(), n, 0);
        mul_vec(va.data(), vb.data(), n);
    }
    transform(va.data(), n, 1);
    mul_num(va.data(), n, {1, 1, 1});
    crt(va.data(), n);
    const size_t len = trim(va, n);
    std::vector<uint64_t> res;
    res.reserve((len * DBITS + 63) / 64);
    u128 t = 0;
    int bits = 0;
    size_t i = 0;
    while (i < len) {
        while (bits < 64 && i < len) {
            t |= (u128)va[i++] << bits;
            bits += DBITS;
        }
        if (bits >= 64) {
            res.push_back((uint64_t)t);
            t >>= 64;
            bits -= 64;
        }
    }
    if (t) res.push_back((uint64_t)t);
    else
        while (!res.empty() && res.back() == 0) res.pop_back();
    return res;
}

// ---- brute route (include/bigint/mul.h:78-104): schoolbook on u64 words, 128-bit products ----

std::vector<uint64_t> brute_mul(const std::vector<uint64_t>& A, const std::vector<uint64_t>& B) {
    std::vector<uint64_t> r(A.size() + B.size(), 0);
    for (size_t i = 0; i < A.size(); ++i) {
        uint64_t c = 0;
        size_t j = 0;
        for (; j < B.size(); ++j) {
            const u128 t = (u128)A[i] * B[j] + r[i + j] + c;
            r[i + j] = (uint64_t)t;
            c = (uint64_t)(t >> 64);
        }
        for (; c; ++j) {
            const u128 t = (u128)r[i + j] + c;
            r[i + j] = (uint64_t)t;
            c = (uint64_t)(t >> 64);
        }
    }
    while (!r.empty() && r.back() == 0) r.pop_back();
    return r;
}

// ---- FFT route (src/mul/fft.cpp:281-386, parameters include/bigint/mul.h:107-187) --------------
// A radix-4 complex FFT in double precision on AVX2, single-threaded like the reference's.  The
// operands are cut into digits of B bits (B chosen by the reference's measured capacity table
// for the total input size), A goes into the real part and B into the imaginary part of one
// transform, the spectrum is squared pointwise, and half the imaginary part of the inverse is the
// digit convolution of A and B (for a square, re = im = A).  Twiddles are generated per block
// and quadrant by a chain of products from exact seeds (the error behaviour the capacity table was
// measured with).

// Largest total input bits (64-bit words x 64) each digit size keeps exact, from the reference's
// worst-case measurement (include/bigint/mul.h:130-152).
constexpr size_t kFftCapacity[21] = {0, 0, 0, 0, 0, 0, 0, 0, 16531472, 11048310, 10485760, 2827308,
                                     1433736, 851968, 799288, 334830, 262144, 128826, 18432, 19456,
                                     18200};
constexpr int kFftMinBits = 10, kFftMaxBits = 20;
constexpr size_t kFftMaxTotalBits = kFftCapacity[kFftMinBits];
constexpr size_t kBruteMaxTotalBits = 2 * 128 * 64;
constexpr size_t kNttMaxTotalBits = ((size_t(1) << 26) - 2) * 28;

int fft_digit_bits(size_t total_bits) {
    for (int b = kFftMaxBits; b >= kFftMinBits; --b)
        if (total_bits <= kFftCapacity[b]) return b;
    return kFftMinBits;
}

size_t pow4_at_least(size_t x) {
    size_t p = std::bit_ceil(x);
    return (std::countr_zero(p) & 1) ? p << 1 : p;
}

struct DVec {  // 32-byte aligned doubles
    double* p = nullptr;
    size_t n = 0;
    explicit DVec(size_t count) : n(count) {
        p = static_cast<double*>(std::aligned_alloc(32, (count * sizeof(double) + 31) & ~size_t(31)));
        std::memset(p, 0, count * sizeof(double));
    }
    ~DVec() { std::free(p); }
    DVec(const DVec&) = delete;
    DVec& operator=(const DVec&) = delete;
};

// in place: element i moves to the base-4 digit reversal j of i (log4 n digits); j is kept as
// a counter that is incremented in reversed base 4 (carry from the top digit downwards)
void digit_reverse4(double* re, double* im, size_t n) {
    const int top = std::countr_zero(n) - 2;  // bit position of the top base-4 digit
    size_t j = 0;
    for (size_t i = 0; i + 1 < n; ++i) {
        if (i < j) {
            std::swap(re[i], re[j]);
            std::swap(im[i], im[j]);
        }
        int sh = top;
        while (((j >> sh) & 3) == 3) {
            j -= size_t(3) << sh;
            sh -= 2;
        }
        j += size_t(1) << sh;
    }
}

inline void cmul4(__m256d ar, __m256d ai, __m256d br, __m256d bi, __m256d& rr, __m256d& ri) {
    rr = _mm256_fmsub_pd(ar, br, _mm256_mul_pd(ai, bi));
    ri = _mm256_fmadd_pd(ar, bi, _mm256_mul_pd(ai, br));
}

// Radix-4 butterfly on four vectors of quadrant values a, b, c, d (b, c, d already twiddled):
// outputs at quadrants 0..3 are (a+c)+(b+d), (a-c)-i(b-d), (a+c)-(b+d), (a-c)+i(b-d) for the
// forward transform (w = exp(-2 pi i / N)); the inverse swaps the second and fourth.
template <bool INV>
inline void bfly4(__m256d& ar, __m256d& ai, __m256d& br, __m256d& bi, __m256d& cr, __m256d& ci,
                  __m256d& dr, __m256d& di) {
    const __m256d s0r = _mm256_add_pd(ar, cr), s0i = _mm256_add_pd(ai, ci);
    const __m256d d0r = _mm256_sub_pd(ar, cr), d0i = _mm256_sub_pd(ai, ci);
    const __m256d s1r = _mm256_add_pd(br, dr), s1i = _mm256_add_pd(bi, di);
    const __m256d d1r = _mm256_sub_pd(br, dr), d1i = _mm256_sub_pd(bi, di);
    ar = _mm256_add_pd(s0r, s1r);
    ai = _mm256_add_pd(s0i, s1i);
    cr = _mm256_sub_pd(s0r, s1r);
    ci = _mm256_sub_pd(s0i, s1i);
    // -i (d1) = (d1i, -d1r)
    const __m256d pr = _mm256_add_pd(d0r, d1i), pi = _mm256_sub_pd(d0i, d1r);  // d0 - i d1
    const __m256d qr = _mm256_sub_pd(d0r, d1i), qi = _mm256_add_pd(d0i, d1r);  // d0 + i d1
    if constexpr (!INV) {
        br = pr, bi = pi, dr = qr, di = qi;
    } else {
        br = qr, bi = qi, dr = pr, di = pi;
    }
}

inline void transpose4(__m256d& r0, __m256d& r1, __m256d& r2, __m256d& r3) {
    const __m256d t0 = _mm256_unpacklo_pd(r0, r1), t1 = _mm256_unpackhi_pd(r0, r1);
    const __m256d t2 = _mm256_unpacklo_pd(r2, r3), t3 = _mm256_unpackhi_pd(r2, r3);
    r0 = _mm256_permute2f128_pd(t0, t2, 0x20);
    r1 = _mm256_permute2f128_pd(t1, t3, 0x20);
    r2 = _mm256_permute2f128_pd(t0, t2, 0x31);
    r3 = _mm256_permute2f128_pd(t1, t3, 0x31);
}

// Per layer (block 4^layer) and quadrant t = 1..3: the step w^{4t}, and exact seeds
// w^{t (m + j)}, j < 4, for every 64th position m.  A chain of twiddle products starts from a
// seed and runs 16 steps: the reference's unbroken chains (up to 16,384 steps at 4^10 points)
// lose up to 3e-13, more than its measured digit-size table leaves room for at 11-14-bit
// digits; re-seeding keeps every product exact at the table's limits.
constexpr size_t kReseed = 64;
struct FftLayer {
    double str[3], sti[3];
    std::vector<double> seed_re, seed_im;  // [q][m / 64][4]
};
const std::vector<FftLayer>& fft_layers() {
    static const std::vector<FftLayer> t = [] {
        std::vector<FftLayer> v(11);
        for (int layer = 2; layer < 11; ++layer) {
            const size_t block = size_t(1) << (2 * layer), sub = block >> 2;
            const size_t nseed = (sub + kReseed - 1) / kReseed;
            const double th = 2.0 * 3.14159265358979323846 / (double)block;
            FftLayer& L = v[layer];
            L.seed_re.resize(3 * nseed * 4);
            L.seed_im.resize(3 * nseed * 4);
            for (int q = 0; q < 3; ++q) {
                for (size_t k = 0; k < nseed; ++k)
                    for (int j = 0; j < 4; ++j) {
                        // the angle's integer part first, so it is exact before the one rounding
                        const double e = (double)(((size_t)(q + 1) * (k * kReseed + j)) % block);
                        L.seed_re[(q * nseed + k) * 4 + j] = std::cos(th * e);
                        L.seed_im[(q * nseed + k) * 4 + j] = -std::sin(th * e);
                    }
                L.str[q] = std::cos(th * (q + 1) * 4);
                L.sti[q] = -std::sin(th * (q + 1) * 4);
            }
        }
        return v;
    }();
    return t;
}

template <bool INV>
void fft4(double* re, double* im, size_t n) {
    digit_reverse4(re, im, n);
    // layer 1: 4-point transforms of consecutive quadruples, four at a time through a transpose
    for (size_t i = 0; i < n; i += 16) {
        __m256d a = _mm256_load_pd(re + i), b = _mm256_load_pd(re + i + 4), c = _mm256_load_pd(re + i + 8),
                d = _mm256_load_pd(re + i + 12);
        __m256d ai = _mm256_load_pd(im + i), bi = _mm256_load_pd(im + i + 4), ci = _mm256_load_pd(im + i + 8),
                di = _mm256_load_pd(im + i + 12);
        transpose4(a, b, c, d);
        transpose4(ai, bi, ci, di);
        bfly4<INV>(a, ai, b, bi, c, ci, d, di);
        transpose4(a, b, c, d);
        transpose4(ai, bi, ci, di);
        _mm256_store_pd(re + i, a), _mm256_store_pd(re + i + 4, b), _mm256_store_pd(re + i + 8, c),
            _mm256_store_pd(re + i + 12, d);
        _mm256_store_pd(im + i, ai), _mm256_store_pd(im + i + 4, bi), _mm256_store_pd(im + i + 8, ci),
            _mm256_store_pd(im + i + 12, di);
    }
    const int layers = std::countr_zero(n) / 2;
    const __m256d neg = _mm256_set1_pd(-0.0);
    for (int layer = 2; layer <= layers; ++layer) {
        const size_t block = size_t(1) << (2 * layer), sub = block >> 2;
        const FftLayer& tl = fft_layers()[layer];
        const size_t nseed = (sub + kReseed - 1) / kReseed;
        __m256d sr[3], si[3];
        for (int q = 0; q < 3; ++q) {
            sr[q] = _mm256_set1_pd(tl.str[q]);
            si[q] = _mm256_set1_pd(INV ? -tl.sti[q] : tl.sti[q]);
        }
        for (size_t i0 = 0; i0 < n; i0 += block) {
            __m256d wr[3], wi[3];
            for (size_t m = 0; m < sub; m += 4) {
                if (m % kReseed == 0) {  // exact seeds for this run of 16 steps
                    for (int q = 0; q < 3; ++q) {
                        const size_t o = ((size_t)q * nseed + m / kReseed) * 4;
                        wr[q] = _mm256_loadu_pd(&tl.seed_re[o]);
                        wi[q] = _mm256_loadu_pd(&tl.seed_im[o]);
                        if constexpr (INV) wi[q] = _mm256_xor_pd(wi[q], neg);
                    }
                }
                double* pa = re + i0 + m;
                double* qa = im + i0 + m;
                __m256d ar = _mm256_load_pd(pa), ai = _mm256_load_pd(qa);
                __m256d br, bi, cr, ci, dr, di;
                cmul4(_mm256_load_pd(pa + sub), _mm256_load_pd(qa + sub), wr[0], wi[0], br, bi);
                cmul4(_mm256_load_pd(pa + 2 * sub), _mm256_load_pd(qa + 2 * sub), wr[1], wi[1], cr, ci);
                cmul4(_mm256_load_pd(pa + 3 * sub), _mm256_load_pd(qa + 3 * sub), wr[2], wi[2], dr, di);
                for (int q = 0; q < 3; ++q) cmul4(wr[q], wi[q], sr[q], si[q], wr[q], wi[q]);
                bfly4<INV>(ar, ai, br, bi, cr, ci, dr, di);
                _mm256_store_pd(pa, ar), _mm256_store_pd(qa, ai);
                _mm256_store_pd(pa + sub, br), _mm256_store_pd(qa + sub, bi);
                _mm256_store_pd(pa + 2 * sub, cr), _mm256_store_pd(qa + 2 * sub, ci);
                _mm256_store_pd(pa + 3 * sub, dr), _mm256_store_pd(qa + 3 * sub, di);
            }
        }
    }
    if constexpr (INV) {
        const __m256d s = _mm256_set1_pd(1.0 / (double)n);
        for (size_t i = 0; i < n; i += 4) {
            _mm256_store_pd(re + i, _mm256_mul_pd(_mm256_load_pd(re + i), s));
            _mm256_store_pd(im + i, _mm256_mul_pd(_mm256_load_pd(im + i), s));
        }
    }
}

// u64 words -> digits of `bits` bits (low digit first) as doubles; returns the digit count
size_t words_to_fft_digits(const std::vector<uint64_t>& w, int bits, double* out) {
    size_t k = 0;
    u128 acc = 0;
    int have = 0;
    const uint64_t mask = (uint64_t(1) << bits) - 1;
    for (uint64_t x : w) {
        acc |= (u128)x << have;
        have += 64;
        while (have >= bits) {
            out[k++] = (double)((uint64_t)acc & mask);
            acc >>= bits;
            have -= bits;
        }
    }
    if (have > 0) out[k++] = (double)((uint64_t)acc & mask);
    return k;
}

std::vector<uint64_t> fft_mul(const std::vector<uint64_t>& A0, const std::vector<uint64_t>& B0, bool same) {
    // low zero words are multiplied away and put back (fft.cpp:284-314)
    size_t za = 0, zb = 0;
    while (za < A0.size() && A0[za] == 0) ++za;
    zb = za;
    if (!same)
        for (zb = 0; zb < B0.size() && B0[zb] == 0;) ++zb;
    if (za == A0.size() || zb == B0.size()) return {};
    const std::vector<uint64_t> A(A0.begin() + za, A0.end());
    const std::vector<uint64_t> B = same ? A : std::vector<uint64_t>(B0.begin() + zb, B0.end());
    const int bits = fft_digit_bits((A.size() + B.size()) * 64);
    const size_t da = (A.size() * 64 + bits - 1) / bits, db = (B.size() * 64 + bits - 1) / bits;
    const size_t n = std::max<size_t>(16, pow4_at_least(da + db - 1));
    DVec re(n), im(n);
    words_to_fft_digits(A, bits, re.p);
    if (same) std::memcpy(im.p, re.p, n * sizeof(double));
    else words_to_fft_digits(B, bits, im.p);
    fft4<false>(re.p, im.p, n);
    for (size_t i = 0; i < n; i += 4) {  // z^2 = (re^2 - im^2) + 2 re im i
        const __m256d r = _mm256_load_pd(re.p + i), m = _mm256_load_pd(im.p + i);
        const __m256d t = _mm256_mul_pd(r, m);
        _mm256_store_pd(re.p + i, _mm256_mul_pd(_mm256_add_pd(r, m), _mm256_sub_pd(r, m)));
        _mm256_store_pd(im.p + i, _mm256_add_pd(t, t));
    }
    fft4<true>(re.p, im.p, n);
    // digits Im / 2, rounded, recombined into words (carries in 128 bits)
    std::vector<uint64_t> res(za + zb, 0);
    size_t top = n;
    while (top > 0 && std::llround(im.p[top - 1]) == 0) --top;
    u128 acc = 0;
    int have = 0;
    for (size_t i = 0; i < top; ++i) {
        acc += (u128)(uint64_t)(std::llround(im.p[i]) / 2) << have;
        have += bits;
        while (have >= 64) {
            res.push_back((uint64_t)acc);
            acc >>= 64;
            have -= 64;
        }
    }
    while (acc) {
        res.push_back((uint64_t)acc);
        acc >>= 64;
    }
    while (!res.empty() && res.back() == 0) res.pop_back();
    return res;
}

// The reference's dispatcher _mul (include/bigint/mul.h:208-220) by total input bits; the SSA
// band above the NTT's limit is outside what the baseline restates (no timed workload reaches it).
enum Route { kBrute = 0, kFft = 1, kNtt = 2, kSsa = 3 };
Route route_of(size_t wa, size_t wb) {
    const size_t total = (wa + wb) * 64;
    if (total <= kBruteMaxTotalBits) return kBrute;
    if (total <= kFftMaxTotalBits) return kFft;
    if (total <= kNttMaxTotalBits) return kNtt;
    return kSsa;
}

std::vector<uint64_t> ref_mul(const std::vector<uint64_t>& A, const std::vector<uint64_t>& B, bool same,
                              bool& ok) {
    ok = true;
    switch (route_of(A.size(), B.size())) {
        case kBrute: return brute_mul(A, same ? A : B);
        case kFft: return fft_mul(A, B, same);
        case kNtt: return ntt_mul(A, B, same);
        default: ok = false; return {};
    }
}

std::vector<uint64_t> digits_to_words(const uint32_t* x, size_t n, size_t off) {
    std::vector<uint64_t> r;
    r.reserve(((n - off) * DBITS + 63) / 64 + 1);
    u128 t = 0;
    size_t bits = 0, i = off;
    while (i < n) {
        while (bits < 64 && i < n) {
            t |= (u128)x[i++] << bits;
            bits += DBITS;
        }
        r.push_back((uint64_t)t);
        t >>= 64;
        bits -= 64;
    }
    if (t) r.push_back((uint64_t)t);
    else
        while (!r.empty() && r.back() == 0) r.pop_back();
    return r;
}

}  // namespace

extern "C" {

void cpu_init_thread_pool(uint32_t n) { pool(n == 0 ? 1 : n); }
uint32_t cpu_pool_threads(void) { return pool().valid() ? pool().threads() : 1u; }

// The route the reference's dispatcher takes for la x lb limbs (0 brute, 1 FFT, 2 NTT, 3 SSA), from
// their u64 word counts after mul_digits' repack (mul.h:245-268).
int cpu_route(size_t la, size_t lb) { return (int)route_of((la * DBITS + 63) / 64, (lb * DBITS + 63) / 64); }

// mul_digits with precision 0 (mul.h:223-318 minus the precision branch) on the reference's
// dispatcher: brute, FFT or NTT by size; out holds la+lb+4.  Returns 2 for a product in the SSA
// band, which is not restated.
int cpu_mul_digits(const uint32_t* a, size_t la, const uint32_t* b, size_t lb, int same,
                   uint32_t* out, size_t* nout) {
    size_t za = 0, zb = 0;
    while (a[za] == 0) ++za;
    if (same) zb = za;
    else
        while (b[zb] == 0) ++zb;
    const std::vector<uint64_t> A = digits_to_words(a, la, za);
    bool ok = true;
    const std::vector<uint64_t> C = same ? ref_mul(A, A, true, ok) : ref_mul(A, digits_to_words(b, lb, zb), false, ok);
    if (!ok) return 2;  // beyond the restated routes
    size_t r = za + zb;
    std::memset(out, 0, r * sizeof(uint32_t));
    u128 t = 0;
    size_t bits = 0, i = 0;
    while (i < C.size()) {
        if (bits < (size_t)DBITS) {
            t |= (u128)C[i++] << bits;
            bits += 64;
        }
        while (bits >= (size_t)DBITS) {
            out[r++] = (uint32_t)t & DMASK;
            t >>= DBITS;
            bits -= DBITS;
        }
    }
    while (t) {
        out[r++] = (uint32_t)t & DMASK;
        t >>= DBITS;
    }
    while (r > za + zb + 1 && out[r - 1] == 0) --r;
    *nout = r;
    return 0;
}

}  // extern "C"
