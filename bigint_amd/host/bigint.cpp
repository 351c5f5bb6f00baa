// Host implementation of BigInt / BigFloat (include/bigint/bigint.h).
//
// Behaviour follows the reference's src/bigint.cpp function by function (cited below) so that
// the same sequence of calls yields the same limbs, point positions and decimal text; the big
// products underneath go to the MI355X engine through mul::mul_digits.
#include "bigint/bigint.h"

#include <algorithm>
#include <atomic>
#include <bit>
#include <cmath>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <map>
#include <mutex>
#include <optional>
#include <thread>
#include <vector>

namespace bigint {

// ---- init_thread_pool (reference bigint.h:35-38, src/mul/ntt_multithread.cpp:287-290, 407-409) --
namespace {
std::once_flag g_pool_once;
std::atomic<uint32_t> g_pool_threads{0};
void pool_set(uint32_t n) {
    std::call_once(g_pool_once, [n] { g_pool_threads.store(n); });
}
}  // namespace

void init_thread_pool(uint32_t n) { pool_set(n == 0 ? 1 : n); }

auto thread_pool_size() -> uint32_t {
    pool_set(std::max(1u, std::thread::hardware_concurrency()));
    return g_pool_threads.load();
}

// ---- limb passes over large operands on worker threads ------------------------------------------
// Additions, subtractions, small-factor products and shifts of operands of millions of limbs
// (the Newton iterations of a 10^8-digit pi) are memory passes; they run in contiguous chunks on
// up to the pool's thread count.  The count is read without fixing the pool: as in the
// reference, only products fix it.
namespace {
constexpr size_t kPassMinLimbs = size_t{1} << 20;  // below this a pass stays on the calling thread
constexpr size_t kPassChunkLimbs = size_t{1} << 18;

size_t pass_threads(size_t n) {
    if (n < kPassMinLimbs) return 1;
    uint32_t p = g_pool_threads.load();
    if (p == 0) p = std::max(1u, std::thread::hardware_concurrency());
    return std::max<size_t>(1, std::min<size_t>(p, n / kPassChunkLimbs));
}

// f(t, lo, hi) for nt contiguous chunks of [0, n); chunk 0 on the calling thread.
template <class F>
void run_chunks(size_t n, size_t nt, F&& f) {
    const size_t per = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (size_t t = 1; t < nt; ++t)
        th.emplace_back([&f, t, per, n] { f(t, std::min(n, t * per), std::min(n, (t + 1) * per)); });
    f(0, 0, std::min(n, per));
    for (auto& x : th) x.join();
}

// A pass whose limbs pass a carry (or borrow) upwards: body(lo, hi) computes limbs [lo, hi) with
// carry-in 0 and returns the carry-out; fix(lo, hi, c) adds carry c into limbs [lo, hi) (a
// ripple from lo) and returns what passes hi.  Chunks run in parallel, then the carries are
// settled chunk by chunk from the bottom.  Returns the carry out of limb n - 1.  The carry into a
// chunk is the sum of its predecessor's own carry-out and what rippled through it; for additions
// and subtractions that sum is at most 1 (a chunk that a carry ripples through entirely had no
// carry-out of its own).
template <class Body, class Fix>
uint128_t carry_pass(size_t n, Body&& body, Fix&& fix) {
    const size_t nt = pass_threads(n);
    if (nt <= 1) return body(size_t{0}, n);
    std::vector<uint128_t> cout(nt);
    run_chunks(n, nt, [&](size_t t, size_t lo, size_t hi) { cout[t] = lo < hi ? body(lo, hi) : 0; });
    const size_t per = (n + nt - 1) / nt;
    uint128_t c = cout[0];
    for (size_t t = 1; t < nt; ++t) {
        const size_t lo = std::min(n, t * per), hi = std::min(n, (t + 1) * per);
        c = (c && lo < hi ? fix(lo, hi, c) : c) + cout[t];
    }
    return c;
}

// d[0, n) *= b in 28-bit limbs; returns the carry out of the top limb.
uint128_t mul_small_pass(uint32_t* d, size_t n, uint64_t b) {
    return carry_pass(
        n,
        [d, b](size_t lo, size_t hi) {
            uint128_t carry = 0;
            for (size_t i = lo; i < hi; ++i) {
                const uint128_t v = (uint128_t)b * d[i] + carry;
                d[i] = (uint32_t)v & DIGIT_MASK;
                carry = v >> DIGIT_BITS;
            }
            return carry;
        },
        [d](size_t lo, size_t hi, uint128_t c) {
            for (size_t i = lo; c && i < hi; ++i) {
                c += d[i];
                d[i] = (uint32_t)c & DIGIT_MASK;
                c >>= DIGIT_BITS;
            }
            return c;
        });
}

// Ripple a carry of 1 (add) or a borrow of 1 (subtract) into r[lo, hi); returns 1 if it passes hi.
uint128_t ripple_add1(uint32_t* r, size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) {
        if (r[i] != DIGIT_MASK) {
            ++r[i];
            return 0;
        }
        r[i] = 0;
    }
    return 1;
}
uint128_t ripple_sub1(uint32_t* r, size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) {
        if (r[i] != 0) {
            --r[i];
            return 0;
        }
        r[i] = DIGIT_MASK;
    }
    return 1;
}
}  // namespace

std::array<size_type, 2> BigInt::DEC_STRING_BRUTE_THRESHOLDS = DEC_STRING_BRUTE_THRESHOLDS_DEFAULT;

// ---- parsing (src/bigint.cpp:25-124) ------------------------------------------------------------
BigInt::BigInt(const std::string& s, bool hex) : BigInt() {
    if (hex) {
        auto val = [](unsigned char c) -> int {
            if (c >= '0' && c <= '9') return c - '0';
            if (c >= 'a' && c <= 'f') return c - 'a' + 10;
            if (c >= 'A' && c <= 'F') return c - 'A' + 10;
            return -1;
        };
        size_type first = 0;
        for (; first < s.size(); ++first) {
            const unsigned char c = (unsigned char)s[first];
            if (c == '-') is_neg_ = true;
            else if (val(c) >= 0) break;
        }
        uint32_t limb = 0, bits = 0;
        for (size_type i = s.size(); i > first; --i) {
            const int v = val((unsigned char)s[i - 1]);
            if (v < 0) continue;
            limb |= (uint32_t)v << bits;
            bits += 4;
            if (bits == DIGIT_BITS) {
                data_.push_back(limb);
                limb = 0;
                bits = 0;
            }
        }
        if (limb) data_.push_back(limb);
    } else {
        size_type first = 0;
        for (; first < s.size(); ++first) {
            const unsigned char c = (unsigned char)s[first];
            if (c >= '0' && c <= '9') break;
            if (c == '-') is_neg_ = true;
        }
        data_ = std::move(convert_from_dec_string(s, first, s.size()).first.data_);
    }
    if (data_.empty()) {
        data_.push_back(0);
        is_neg_ = false;
    } else {
        remove_leading_zero();
    }
}

auto BigInt::convert_from_dec_string(const std::string& s, size_type start, size_type end)
    -> std::pair<BigInt, size_type> {
    if (end - start < std::max<size_type>(2, DEC_STRING_BRUTE_THRESHOLDS[0])) {
        BigInt res(0);
        size_type ndig = 0;
        uint64_t chunk = 0, scale = 1;
        for (size_type i = start; i < end; ++i) {
            const unsigned char c = (unsigned char)s[i];
            if (c < '0' || c > '9') continue;
            ++ndig;
            scale *= TEN;
            chunk = chunk * TEN + (c - '0');
            if (scale > UINT64_MAX / TEN) {
                res *= scale;
                res += chunk;
                scale = 1;
                chunk = 0;
            }
        }
        if (scale > 1) {
            res *= scale;
            res += chunk;
        }
        return {std::move(res), ndig};
    }
    const size_type mid = (start + end) / 2;
    auto hi = convert_from_dec_string(s, start, mid);
    auto lo = convert_from_dec_string(s, mid, end);
    return {hi.first * get_pow_of_ten((uint32_t)lo.second) + lo.first, hi.second + lo.second};
}

// ---- printing (src/bigint.cpp:126-265) ------------------------------------------------------------
auto BigInt::to_dec_string_brute() const -> std::string {
    if (is_zero()) return "0";
    // Repeated division of the magnitude by 10^10, ten digits per step: every partial dividend
    // rem * 2^28 + limb is below 10^10 * 2^28 < 2^62, so each limb costs one 64-bit division by a
    // constant (a multiply-high), and each quotient limb stays below 2^28.
    constexpr uint64_t chunk = 10000000000ull;
    constexpr int chunk_digits = 10;
    std::vector<uint32_t> t(data_.begin(), data_.end());
    size_type n = t.size();
    std::string r;
    r.reserve(chunk_digits + 1 + (size_type)std::ceil((double)data_.size() * DIGIT_BITS * std::log10(2.0)));
    while (n > 0) {
        uint64_t rem = 0;
        for (size_type i = n; i-- > 0;) {
            const uint64_t cur = (rem << DIGIT_BITS) | t[i];
            const uint64_t q = cur / chunk;
            t[i] = (uint32_t)q;
            rem = cur - q * chunk;
        }
        while (n > 0 && t[n - 1] == 0) --n;
        for (int i = 0; i < chunk_digits && (rem != 0 || n > 0); ++i) {
            r.push_back((char)('0' + rem % TEN));
            rem /= TEN;
        }
    }
    if (is_neg_) r.push_back('-');
    std::reverse(r.begin(), r.end());
    return r;
}

auto BigInt::get_pow_of_ten(uint32_t exponent) -> BigInt {
    static std::vector<BigInt> squares = {BigInt(TEN)};  // 10^(2^i)
    static std::mutex mtx;
    std::lock_guard<std::mutex> lk(mtx);
    BigInt res(1);
    if (exponent == 0) return res;
    const uint32_t nb = (uint32_t)std::bit_width(exponent);
    while (squares.size() < nb) squares.push_back(squares.back() * squares.back());
    for (uint32_t i = 0; i < nb; ++i)
        if ((exponent >> i) & 1u) res *= squares[i];
    return res;
}

void BigInt::print(std::ostream& output, bool hex, bool direct) const {
    std::optional<std::stringstream> buf;
    std::ostream& out = direct ? output : buf.emplace();
    if (hex) print_hex(out);
    else print_dec(out);
    if (!direct) output << buf->rdbuf();
}

// Same digits as the reference's recursive split (src/bigint.cpp:126-265): x = q 10^k + r with
// k half of x's decimal length, down to pieces below DEC_STRING_BRUTE_THRESHOLDS[1] limbs.
// The decimal string of a number is unique, so five things here only change the speed:
//   - the split tree is walked one depth at a time: the quotient estimates q = floor(x / 10^k) of
//     all nodes of a depth are one product batch (x times a reciprocal of 10^k), and so are the
//     products q 10^k; the estimates are then corrected until exact;
//   - 10^k and its reciprocal are computed once per distinct k (a depth has one or two), the
//     reciprocal only as precise as the quotient needs (about half of x's limbs);
//   - subtractions, corrections and the brute-force conversions of the pieces run on the
//     init_thread_pool threads;
//   - digits are written in blocks.
void BigInt::print_dec(std::ostream& output, size_type len) const {
    if (is_neg_) {  // the sign, then the magnitude padded to len digits
        output.put('-');
        BigInt m(*this);
        m.is_neg_ = false;
        m.print_dec(output, len);
        return;
    }
    const size_type thr = DEC_STRING_BRUTE_THRESHOLDS[1];
    auto small = [&](const BigInt& x) { return x.data_.size() == 1 || x.data_.size() < thr; };
    auto parallel_for = [](size_t count, auto&& f) {
        const size_t nt = std::min<size_t>(thread_pool_size(), count);
        std::atomic<size_t> next{0};
        auto work = [&] {
            for (size_t i; (i = next.fetch_add(1)) < count;) f(i);
        };
        std::vector<std::thread> threads;
        for (size_t t = 1; t < nt; ++t) threads.emplace_back(work);
        work();
        for (auto& t : threads) t.join();
    };
    struct Div {
        BigInt d;      // 10^k
        BigInt recip;  // digits of a reciprocal of 10^k: 1/10^k ~ recip * 2^(-28 shift)
        int64_t shift;
        size_type prec;
        BigFloat rc;   // the reciprocal itself
    };
    std::map<size_type, Div> divs;
    std::vector<std::pair<BigInt, size_type>> pieces;  // in output order: (value, width)
    pieces.emplace_back(*this, len);
    for (;;) {
        std::vector<size_t> at;  // pieces split at this depth
        for (size_t i = 0; i < pieces.size(); ++i)
            if (!small(pieces[i].first)) at.push_back(i);
        if (at.empty()) break;
        std::vector<size_type> ks(at.size());
        std::map<size_type, size_type> need;  // k -> precision in limbs
        for (size_t j = 0; j < at.size(); ++j) {
            const BigInt& x = pieces[at[j]].first;
            const size_type est = (size_type)std::ceil((double)x.data_.size() * DIGIT_BITS * std::log10(2.0));
            ks[j] = est / 2;
            // q = floor(x / 10^k) has about |x| - |10^k| limbs; a reciprocal with relative error
            // below 2^(-28 p) puts the estimate x * recip within 1 of x / 10^k once p exceeds
            // |q| + 1 limbs, so p = |x| - (limbs of 10^k, rounded down) + 8 leaves a margin
            const size_type dlimbs = (size_type)((double)ks[j] * std::log2(10.0) / DIGIT_BITS);
            const size_type prec = x.data_.size() + 8 - std::min(dlimbs, x.data_.size());
            size_type& p = need[ks[j]];
            p = std::max(p, prec);
        }
        for (const auto& [k, prec] : need) {
            auto it = divs.find(k);
            if (it != divs.end() && it->second.prec >= prec + 2) continue;
            BigInt d = it == divs.end() ? get_pow_of_ten((uint32_t)k) : std::move(it->second.d);
            // a reciprocal of a larger power at this precision or better, from a shallower depth,
            // gives this one by one product: 1/10^k = 1/10^k2 * 10^(k2 - k) (relative error of the
            // parent's, plus the truncation to prec + 2 limbs); else Newton's iteration
            const Div* parent = nullptr;
            size_type k2 = 0;
            for (auto up = divs.upper_bound(k); up != divs.end(); ++up)
                if (up->second.prec >= prec + 2) {
                    parent = &up->second;
                    k2 = up->first;
                    break;
                }
            BigFloat rc;
            if (parent) {
                // 10^(k2 - k) from 10^k: k2 - k is within a few units of k (a child holds about
                // half of its parent's digits)
                const int64_t dm = (int64_t)(k2 - k) - (int64_t)k;
                uint64_t p10 = 1;
                for (int64_t i = 0; i < (dm < 0 ? -dm : dm) && i < 19; ++i) p10 *= 10u;
                BigInt pw = dm >= 0 && dm <= 18 ? d * p10
                            : dm < 0 && dm >= -18 ? d / p10
                                                  : get_pow_of_ten((uint32_t)(k2 - k));
                rc = BigFloat::mul(parent->rc, BigFloat(std::move(pw)), prec + 2);
            } else {
                rc = BigFloat(d).reciprocal(prec + 2);
            }
            BigInt r;
            r.data_ = rc.get_data();
            r.remove_leading_zero();
            const int64_t shift = rc.get_point_pos();
            divs[k] = Div{std::move(d), std::move(r), shift, prec + 2, std::move(rc)};
        }
        // q = floor(x * recip * 2^(-28 shift)): within a few units of floor(x / 10^k)
        std::vector<std::pair<const BigInt*, const BigInt*>> jobs(at.size());
        for (size_t j = 0; j < at.size(); ++j) jobs[j] = {&pieces[at[j]].first, &divs.at(ks[j]).recip};
        std::vector<BigInt> qs = mul_batch(jobs);
        parallel_for(at.size(), [&](size_t j) { qs[j] >>= (int64_t)DIGIT_BITS * divs.at(ks[j]).shift; });
        for (size_t j = 0; j < at.size(); ++j) jobs[j] = {&qs[j], &divs.at(ks[j]).d};
        std::vector<BigInt> rs = mul_batch(jobs);
        parallel_for(at.size(), [&](size_t j) {
            const BigInt& d = divs.at(ks[j]).d;
            BigInt& q = qs[j];
            BigInt& r = rs[j];
            r = pieces[at[j]].first - r;
            while (r.is_neg_) {
                r += d;
                --q;
            }
            while (compare_abs(r, d) >= 0) {
                r -= d;
                ++q;
            }
        });
        std::vector<std::pair<BigInt, size_type>> next;
        next.reserve(pieces.size() + at.size());
        for (size_t i = 0, j = 0; i < pieces.size(); ++i) {
            if (j < at.size() && at[j] == i) {
                const size_type width = pieces[i].second, k = ks[j];
                next.emplace_back(std::move(qs[j]), width > k ? width - k : 0);
                next.emplace_back(std::move(rs[j]), k);
                ++j;
            } else {
                next.push_back(std::move(pieces[i]));
            }
        }
        pieces = std::move(next);
    }
    std::vector<std::string> strs(pieces.size());
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t i; (i = next.fetch_add(1)) < pieces.size();) strs[i] = pieces[i].first.to_dec_string_brute();
    };
    const size_t nt = std::min<size_t>(thread_pool_size(), pieces.size());
    std::vector<std::thread> threads;
    for (size_t t = 1; t < nt; ++t) threads.emplace_back(work);
    work();
    for (auto& t : threads) t.join();
    const std::string zeros(4096, '0');
    for (size_t i = 0; i < pieces.size(); ++i) {
        const std::string& str = strs[i];
        const size_type width = pieces[i].second;
        size_type body = str.size(), start = 0;
        if (pieces[i].first.is_neg_) {
            output.put(str[0]);
            --body;
            start = 1;
        }
        for (size_type pad = body < width ? width - body : 0; pad;) {
            const size_type c = std::min<size_type>(pad, zeros.size());
            output.write(zeros.data(), (std::streamsize)c);
            pad -= c;
        }
        output.write(str.data() + start, (std::streamsize)body);
    }
}

void BigInt::print_hex(std::ostream& output) const {
    if (is_zero()) {
        output.put('0');
        return;
    }
    if (is_neg_) output.put('-');
    output << std::hex << data_.back() << std::setfill('0');
    for (size_type i = data_.size() - 1; i-- > 0;) output << std::setw(DIGIT_BITS / 4) << data_[i];
    output << std::dec << std::setfill(' ');
}

auto operator<<(std::ostream& output, const BigInt& a) -> std::ostream& {
    a.print(output, (output.flags() & std::ios_base::hex) != 0, false);
    return output;
}

auto operator>>(std::istream& input, BigInt& a) -> std::istream& {
    std::string s;
    input >> s;
    a = BigInt(s, (input.flags() & std::ios_base::hex) != 0);
    return input;
}

// ---- basic magnitude arithmetic (src/bigint.cpp:267-548) ------------------------------------------
auto BigInt::remove_leading_zero() -> size_type {
    size_type n = data_.size();
    while (n > 1 && data_[n - 1] == 0) --n;
    if (n == 1 && data_[0] == 0) {
        reset();
        return 1;
    }
    data_.resize(n);
    return n;
}

auto BigInt::compare_abs(const BigInt& a, const BigInt& b) noexcept -> std::strong_ordering {
    if (auto c = a.data_.size() <=> b.data_.size(); c != 0) return c;
    for (size_type i = a.data_.size(); i-- > 0;)
        if (auto c = a.data_[i] <=> b.data_[i]; c != 0) return c;
    return std::strong_ordering::equal;
}

void BigInt::unsigned_self_inc_or_dec(bool is_inc) {
    size_type i = 0;
    const size_type n = data_.size();
    if (is_inc) {
        while (i < n && data_[i] == DIGIT_MASK) data_[i++] = 0;
        if (i == n) data_.push_back(1);
        else ++data_[i];
    } else {
        while (i < n && data_[i] == 0) data_[i++] = DIGIT_MASK;
        if (i < n) --data_[i];
    }
}

auto BigInt::operator++() -> BigInt& {
    if (is_neg_) {
        unsigned_self_inc_or_dec(false);
        remove_leading_zero();
    } else {
        unsigned_self_inc_or_dec(true);
    }
    return *this;
}
auto BigInt::operator++(int) -> BigInt {
    BigInt t = *this;
    ++*this;
    return t;
}
auto BigInt::operator--() -> BigInt& {
    if (is_zero()) {
        data_[0] = 1;
        is_neg_ = true;
    } else if (is_neg_) {
        unsigned_self_inc_or_dec(true);
    } else {
        unsigned_self_inc_or_dec(false);
        remove_leading_zero();
    }
    return *this;
}
auto BigInt::operator--(int) -> BigInt {
    BigInt t = *this;
    --*this;
    return t;
}

namespace {
// |x| + |y| into r (limbs), canonical
void mag_add(Digits& r, const Digits& x, const Digits& y) {
    const size_type n = std::max(x.size(), y.size());
    const size_type nx = x.size(), ny = y.size();
    r.resize(n);  // r may be x or y: only limbs past the old end are new
    const uint32_t* xp = x.data();
    const uint32_t* yp = y.data();
    uint32_t* rp = r.data();
    const uint128_t carry = carry_pass(
        n,
        [=](size_t lo, size_t hi) {
            uint32_t c = 0;
            for (size_t i = lo; i < hi; ++i) {
                const uint32_t v = (i < nx ? xp[i] : 0u) + (i < ny ? yp[i] : 0u) + c;
                rp[i] = v & DIGIT_MASK;
                c = v >> DIGIT_BITS;
            }
            return (uint128_t)c;
        },
        [=](size_t lo, size_t hi, uint128_t) { return ripple_add1(rp, lo, hi); });
    if (carry) r.push_back((uint32_t)carry);
}
// |x| - |y| into r with |x| >= |y|
void mag_sub(Digits& r, const Digits& x, const Digits& y) {
    const size_type nx = x.size(), ny = y.size();
    r.resize(nx);
    const uint32_t* xp = x.data();
    const uint32_t* yp = y.data();
    uint32_t* rp = r.data();
    (void)carry_pass(
        nx,
        [=](size_t lo, size_t hi) {
            uint32_t borrow = 0;
            for (size_t i = lo; i < hi; ++i) {
                const uint32_t v = xp[i] - (i < ny ? yp[i] : 0u) - borrow;
                rp[i] = v & DIGIT_MASK;
                borrow = v >> 31;  // wrapped
            }
            return (uint128_t)borrow;
        },
        [=](size_t lo, size_t hi, uint128_t) { return ripple_sub1(rp, lo, hi); });
}
}  // namespace

void BigInt::inplace_add_or_sub(const BigInt& b, bool is_sub) {
    if (is_neg_ == (b.is_neg_ != is_sub)) {  // same effective sign: magnitudes add
        mag_add(data_, data_, b.data_);
        return;
    }
    const auto c = compare_abs(*this, b);
    if (c == 0) {
        reset();
        return;
    }
    if (c > 0) {
        mag_sub(data_, data_, b.data_);
    } else {
        Digits t;
        mag_sub(t, b.data_, data_);
        data_ = std::move(t);
        is_neg_ = !is_neg_;
    }
    remove_leading_zero();
}

void BigInt::inplace_add_or_sub(uint64_t b, bool is_sub) {
    if (b == 0) return;
    BigInt t(b, is_sub);
    inplace_add_or_sub(t, false);
}

void BigInt::unsigned_inplace_mul(uint64_t b) {
    if (is_zero() || b == 1) return;
    if (b == 0) {
        reset();
        return;
    }
    if ((b & (b - 1)) == 0) {
        *this <<= std::countr_zero(b);
        return;
    }
    uint128_t carry = mul_small_pass(data_.data(), data_.size(), b);
    while (carry) {
        data_.push_back((uint32_t)carry & DIGIT_MASK);
        carry >>= DIGIT_BITS;
    }
    remove_leading_zero();
}

auto BigInt::unsigned_inplace_divmod(uint64_t b) -> uint64_t {
    if (b == 0) throw std::domain_error("division by zero");
    uint128_t rem = 0;
    for (size_type i = data_.size(); i-- > 0;) {
        rem = (rem << DIGIT_BITS) + data_[i];
        data_[i] = (uint32_t)(rem / b);
        rem %= b;
    }
    const bool neg = is_neg_;
    remove_leading_zero();
    if (!is_zero()) is_neg_ = neg;
    return (uint64_t)rem;
}

void BigInt::shift_left(Digits& v, uint64_t offset) {
    const uint64_t whole = offset / DIGIT_BITS, bits = offset % DIGIT_BITS;
    if (bits == 0) {
        v.insert(v.begin(), whole, 0u);
        return;
    }
    const size_type n = v.size();
    v.resize(n + whole + 1);
    // high to low: new[i] = (old[i-whole] << bits | old[i-whole-1] >> (28-bits)) & mask
    for (size_type i = n + whole; i > whole; --i) {
        const uint32_t lo_src = v[i - whole - 1];
        const uint32_t hi_src = (i - whole < n) ? v[i - whole] : 0u;
        v[i] = ((hi_src << bits) | (lo_src >> (DIGIT_BITS - bits))) & DIGIT_MASK;
    }
    v[whole] = (v[0] << bits) & DIGIT_MASK;
    std::fill(v.begin(), v.begin() + (std::ptrdiff_t)whole, 0u);
}

void BigInt::shift_right(Digits& v, uint64_t offset) {
    const uint64_t whole = offset / DIGIT_BITS, bits = offset % DIGIT_BITS;
    if (whole >= v.size()) {
        v.assign(1, 0u);
        return;
    }
    if (bits == 0) {
        v.erase(v.begin(), v.begin() + (std::ptrdiff_t)whole);
        return;
    }
    const size_type n = v.size() - whole;
    for (size_type i = 0; i + 1 < n; ++i)
        v[i] = ((v[i + whole] >> bits) | (v[i + whole + 1] << (DIGIT_BITS - bits))) & DIGIT_MASK;
    v[n - 1] = v[v.size() - 1] >> bits;
    v.resize(n);
}

// ---- bitwise (src/bigint.cpp:551-623) -------------------------------------------------------------
auto BigInt::operator&=(const BigInt& b) -> BigInt& {
    if (is_zero()) return *this;
    if (b.is_zero()) {
        reset();
        return *this;
    }
    if (data_.size() > b.data_.size()) data_.resize(b.data_.size());
    for (size_type i = 0; i < data_.size(); ++i) data_[i] &= b.data_[i];
    remove_leading_zero();
    return *this;
}
auto BigInt::operator|=(const BigInt& b) -> BigInt& {
    if (is_zero()) return *this = b;
    if (b.is_zero()) return *this;
    const size_type n = data_.size();
    if (n < b.data_.size()) data_.insert(data_.end(), b.data_.begin() + (std::ptrdiff_t)n, b.data_.end());
    for (size_type i = 0; i < std::min(n, b.data_.size()); ++i) data_[i] |= b.data_[i];
    remove_leading_zero();
    return *this;
}
auto BigInt::operator^=(const BigInt& b) -> BigInt& {
    if (is_zero()) return *this = b;
    if (b.is_zero()) return *this;
    const size_type n = data_.size();
    if (n < b.data_.size()) data_.insert(data_.end(), b.data_.begin() + (std::ptrdiff_t)n, b.data_.end());
    for (size_type i = 0; i < std::min(n, b.data_.size()); ++i) data_[i] ^= b.data_[i];
    remove_leading_zero();
    return *this;
}
auto BigInt::bitwise_not(size_type len) -> BigInt& {
    if (len) data_.resize(len);
    for (auto& d : data_) d = ~d & DIGIT_MASK;
    remove_leading_zero();
    return *this;
}

auto BigInt::from_digits(Digits digits, bool negative) -> BigInt {
    BigInt r;
    r.data_ = std::move(digits);
    if (r.data_.empty()) r.data_.push_back(0);
    r.is_neg_ = negative;
    r.remove_leading_zero();
    return r;
}

auto BigInt::mul_batch(const std::vector<std::pair<const BigInt*, const BigInt*>>& pairs,
                       int plan_mode) -> std::vector<BigInt> {
    std::vector<std::pair<const Digits*, const Digits*>> dp;
    dp.reserve(pairs.size());
    for (const auto& [pa, pb] : pairs)
        if (!pa->is_zero() && !pb->is_zero()) dp.emplace_back(&pa->data_, &pb->data_);
    auto prods = mul::mul_digits_batch(dp, plan_mode);
    std::vector<BigInt> res;
    res.reserve(pairs.size());
    size_t k = 0;
    for (const auto& [pa, pb] : pairs) {
        if (pa->is_zero() || pb->is_zero()) {
            res.push_back(BigInt(0));
            continue;
        }
        BigInt r;
        r.data_ = std::move(prods[k++]);
        r.is_neg_ = pa->is_neg_ != pb->is_neg_;
        r.remove_leading_zero();
        res.push_back(std::move(r));
    }
    return res;
}

// ---- division via the reciprocal (src/bigint.cpp:625-677) -----------------------------------------
auto BigInt::divmod(const BigInt& b, RoundMode mode, bool check) const -> std::pair<BigInt, BigInt> {
    if (b.is_zero()) throw std::domain_error("division by zero");
    if (is_zero()) return {BigInt(0), BigInt(0)};
    constexpr size_type guard = 16;
    const BigInt& a = *this;
    const size_type prec = a.data_.size() + guard;
    BigInt q(BigFloat::mul(BigFloat(a), BigFloat(b).reciprocal(prec), prec), mode);
    BigInt r(a - q * b);
    if (!check || r.is_zero()) return {std::move(q), std::move(r)};
    bool fix = false;
    switch (mode) {
        case RoundMode::Floor: fix = (r.is_neg_ != b.is_neg_) || compare_abs(r, b) >= 0; break;
        case RoundMode::Ceil: fix = (r.is_neg_ == b.is_neg_) || compare_abs(r, b) >= 0; break;
        case RoundMode::RoundHalfUp: {
            const auto c = compare_abs(r * 2, b);
            fix = c > 0 || (c == 0 && r.is_neg_ == a.is_neg_);
            break;
        }
        case RoundMode::Truncate:
        default: fix = (r.is_neg_ != a.is_neg_) || compare_abs(r, b) >= 0; break;
    }
    if (fix) {
        if (r.is_neg_ != b.is_neg_) {
            r += b;
            --q;
        } else {
            r -= b;
            ++q;
        }
    }
    return {std::move(q), std::move(r)};
}

// ---- BigFloat ------------------------------------------------------------------------------------
BigFloat::BigFloat(const BigInt& x, int64_t offset) : data_(x.data_), point_pos_(0), is_neg_(x.is_neg_) {
    *this <<= offset;
    remove_tail_zero();
}

BigFloat::BigFloat(BigInt&& x, int64_t offset)
    : data_(std::move(x.data_)), point_pos_(0), is_neg_(x.is_neg_) {
    x.reset();
    *this <<= offset;
    remove_tail_zero();
}

// src/bigint.cpp:681-715
BigFloat::BigFloat(double value) : point_pos_(0), is_neg_(false) {
    if (std::isinf(value)) throw std::domain_error("invalid double to BigFloat conversion: inf");
    if (std::isnan(value)) throw std::domain_error("invalid double to BigFloat conversion: nan");
    if (value == 0) {
        data_.push_back(0);
        return;
    }
    const uint64_t raw = std::bit_cast<uint64_t>(value);
    is_neg_ = (raw >> 63) != 0;
    const int64_t ebits = (int64_t)((raw >> DOUBLE_MANTISSA_LEN) & ((1u << DOUBLE_EXPONENT_LEN) - 1));
    uint64_t mant = raw & ((1ull << DOUBLE_MANTISSA_LEN) - 1);
    int64_t offset = -DOUBLE_MANTISSA_LEN - DOUBLE_EXPONENT_BIAS;
    if (ebits == 0) {
        offset += 1;
    } else {
        mant |= 1ull << DOUBLE_MANTISSA_LEN;
        offset += ebits;
    }
    while (mant) {
        data_.push_back((uint32_t)mant & DIGIT_MASK);
        mant >>= DIGIT_BITS;
    }
    shift(offset);
}

// src/bigint.cpp:717-766
auto BigFloat::to_double() const -> double {
    if (is_zero()) return 0;
    const int64_t hw = std::bit_width(data_.back());
    const int64_t exponent = (int64_t)DIGIT_BITS * ((int64_t)data_.size() - 1 - point_pos_) + hw - 1;
    int64_t ebits = exponent + DOUBLE_EXPONENT_BIAS;
    if (ebits >= (1 << DOUBLE_EXPONENT_LEN) - 1) return is_neg_ ? -HUGE_VAL : HUGE_VAL;
    auto shl = [](uint64_t x, int64_t r) -> uint64_t {
        if (r <= -64 || r >= 64) return 0;
        return r < 0 ? x >> (-r) : x << r;
    };
    int64_t off;
    uint64_t mbits;
    if (ebits <= 0) {
        constexpr int64_t min_sub = -DOUBLE_MANTISSA_LEN - DOUBLE_EXPONENT_BIAS + 1;
        if (exponent < min_sub) return is_neg_ ? -0.0 : 0.0;
        ebits = 0;
        off = exponent + DOUBLE_MANTISSA_LEN + DOUBLE_EXPONENT_BIAS - hw;
        mbits = shl(data_.back(), off);
    } else {
        off = DOUBLE_MANTISSA_LEN - (hw - 1);
        mbits = shl(data_.back() & ((1u << (hw - 1)) - 1), off);
    }
    size_type i = data_.size() - 1;
    while (off > 0 && i > 0) {
        off -= DIGIT_BITS;
        --i;
        mbits |= shl(data_[i], off);
    }
    return std::bit_cast<double>(((uint64_t)is_neg_ << 63) | ((uint64_t)ebits << DOUBLE_MANTISSA_LEN) | mbits);
}

// src/bigint.cpp:769-812
void BigFloat::shift(int64_t offset) {
    point_pos_ -= offset / (int64_t)DIGIT_BITS;
    offset %= (int64_t)DIGIT_BITS;
    if (offset == 0) {
        remove_leading_zero();
        return;
    }
    if (offset > 0) {
        offset = DIGIT_BITS - offset;
        --point_pos_;
    } else {
        offset = -offset;
    }
    if (std::countr_zero(data_[0]) >= offset) {
        shift_right(data_, (uint32_t)offset);
    } else {
        shift_left(data_, DIGIT_BITS - (uint32_t)offset);
        ++point_pos_;
    }
    remove_leading_zero();
}

// In place, in chunks: each chunk reads the neighbour limb across its edge before any chunk
// writes (edge[]), so the chunks are independent.
void BigFloat::shift_right(Digits& v, uint32_t offset) {
    const size_t n = v.size();
    uint32_t* d = v.data();
    const size_t nt = pass_threads(n), per = (n + nt - 1) / nt;
    if (nt == 1) {
        for (size_t i = 0; i + 1 < n; ++i) d[i] = ((d[i] >> offset) | (d[i + 1] << (DIGIT_BITS - offset))) & DIGIT_MASK;
        d[n - 1] >>= offset;
        return;
    }
    std::vector<uint32_t> edge(nt);  // limb above each chunk (0 above the top)
    for (size_t t = 0; t < nt; ++t) edge[t] = (t + 1) * per < n ? d[(t + 1) * per] : 0u;
    run_chunks(n, nt, [&](size_t t, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            const uint32_t above = i + 1 < hi ? d[i + 1] : edge[t];
            d[i] = ((d[i] >> offset) | (above << (DIGIT_BITS - offset))) & DIGIT_MASK;
        }
    });
}

void BigFloat::shift_left(Digits& v, uint32_t offset) {
    v.push_back(0);
    const size_t n = v.size();
    uint32_t* d = v.data();
    const size_t nt = pass_threads(n), per = (n + nt - 1) / nt;
    if (nt == 1) {
        for (size_t i = n - 1; i > 0; --i) d[i] = ((d[i] << offset) | (d[i - 1] >> (DIGIT_BITS - offset))) & DIGIT_MASK;
        d[0] = (d[0] << offset) & DIGIT_MASK;
        return;
    }
    std::vector<uint32_t> edge(nt);  // limb below each chunk (0 below the bottom)
    for (size_t t = 0; t < nt; ++t) edge[t] = t > 0 && t * per <= n ? d[t * per - 1] : 0u;
    run_chunks(n, nt, [&](size_t t, size_t lo, size_t hi) {
        for (size_t i = hi; i-- > lo;) {
            const uint32_t below = i > lo ? d[i - 1] : edge[t];
            d[i] = ((d[i] << offset) | (below >> (DIGIT_BITS - offset))) & DIGIT_MASK;
        }
    });
}

auto BigFloat::remove_leading_zero() -> size_type {
    size_type n = data_.size();
    while (n > 1 && data_[n - 1] == 0) --n;
    if (n == 1 && data_[0] == 0) {
        reset();
        return 1;
    }
    data_.resize(n);
    return n;
}

auto BigFloat::remove_tail_zero() -> size_type {
    if (is_zero()) return 1;
    size_type z = 0;
    while (data_[z] == 0) ++z;
    if (z) {
        point_pos_ -= (int64_t)z;
        data_.erase(data_.begin(), data_.begin() + (std::ptrdiff_t)z);
    }
    return data_.size();
}

// src/bigint.cpp:849-936
auto BigFloat::add_or_sub(const BigFloat& a, const BigFloat& b, bool is_sub) -> BigFloat {
    const int64_t la = (int64_t)a.data_.size(), lb = (int64_t)b.data_.size();
    int64_t tza = 0, tzb = 0;
    while (tza < la && a.data_[tza] == 0) ++tza;
    while (tzb < lb && b.data_[tzb] == 0) ++tzb;
    if (tza == la) return is_sub ? -b : b;
    if (tzb == lb) return a;
    const int64_t pp = std::max(a.point_pos_ - tza, b.point_pos_ - tzb);
    const int64_t lr = std::max(la - a.point_pos_, lb - b.point_pos_) + pp;
    const int64_t oa = pp - a.point_pos_, ob = pp - b.point_pos_;
    // limb i of the result frame holds a.data_[i - oa] for i in [a_lo, a_hi), likewise for b
    const int64_t a_lo = tza + oa, a_hi = la + oa, b_lo = tzb + ob, b_hi = lb + ob;
    const uint32_t* ad = a.data_.data() - oa;
    const uint32_t* bd = b.data_.data() - ob;
    auto A = [&](int64_t i) -> uint32_t { return i >= a_lo && i < a_hi ? ad[i] : 0u; };
    auto B = [&](int64_t i) -> uint32_t { return i >= b_lo && i < b_hi ? bd[i] : 0u; };
    const bool subtract = a.is_neg_ != (b.is_neg_ != is_sub);
    BigFloat res;
    res.point_pos_ = pp;
    res.data_.resize((size_t)lr + 1);
    uint32_t* r = res.data_.data();
    if (!subtract) {  // |a| + |b| with the sign of a
        res.is_neg_ = a.is_neg_;
        const uint128_t c = carry_pass(
            (size_t)lr,
            [&](size_t lo, size_t hi) {
                uint32_t cc = 0;
                for (int64_t i = (int64_t)lo; i < (int64_t)hi; ++i) {
                    const uint32_t t = A(i) + B(i) + cc;
                    r[i] = t & DIGIT_MASK;
                    cc = t >> DIGIT_BITS;
                }
                return (uint128_t)cc;
            },
            [r](size_t lo, size_t hi, uint128_t) { return ripple_add1(r, lo, hi); });
        r[lr] = (uint32_t)c;
    } else {  // the larger magnitude minus the smaller, with the larger one's sign
        int64_t top = lr - 1;
        while (top >= 0 && A(top) == B(top)) --top;
        if (top < 0) {
            res.data_.assign(1, 0u);
            res.point_pos_ = 0;
            res.is_neg_ = false;
            return res;
        }
        const bool a_big = A(top) > B(top);
        res.is_neg_ = a_big ? a.is_neg_ : !a.is_neg_;
        (void)carry_pass(
            (size_t)lr,
            [&](size_t lo, size_t hi) {
                uint32_t borrow = 0;
                for (int64_t i = (int64_t)lo; i < (int64_t)hi; ++i) {
                    const uint32_t x = a_big ? A(i) : B(i);
                    const uint32_t y = (a_big ? B(i) : A(i)) + borrow;
                    borrow = x < y;
                    r[i] = (x + (borrow << DIGIT_BITS) - y) & DIGIT_MASK;
                }
                return (uint128_t)borrow;
            },
            [r](size_t lo, size_t hi, uint128_t) { return ripple_sub1(r, lo, hi); });
        r[lr] = 0;
    }
    res.remove_leading_zero();
    res.remove_tail_zero();
    return res;
}

void BigFloat::unsigned_inplace_mul(uint64_t b) {
    if (is_zero() || b == 1) return;
    if (b == 0) {
        reset();
        return;
    }
    if ((b & (b - 1)) == 0) {
        *this <<= std::countr_zero(b);
        return;
    }
    uint128_t carry = mul_small_pass(data_.data(), data_.size(), b);
    while (carry) {
        data_.push_back((uint32_t)carry & DIGIT_MASK);
        carry >>= DIGIT_BITS;
    }
    remove_leading_zero();
}

// src/bigint.cpp:972-1042
void BigFloat::print(std::ostream& output, size_type dec_digits, bool direct) const {
    std::optional<std::stringstream> buf;
    std::ostream& out = direct ? output : buf.emplace();
    BigInt ip, fp;
    if (point_pos_ < (int64_t)data_.size()) {
        if (point_pos_ < 0) ip.data_.assign((size_type)(-point_pos_), 0u);
        ip.data_.insert(ip.data_.end(), data_.begin() + std::max<int64_t>(0, point_pos_), data_.end());
    }
    if (ip.data_.empty()) ip.data_.push_back(0);
    if (point_pos_ > 0)
        fp.data_.assign(data_.begin(), data_.begin() + std::min<int64_t>(point_pos_, (int64_t)data_.size()));
    if (dec_digits == 0)
        dec_digits = (size_type)std::floor((double)(fp.data_.size() * DIGIT_BITS) * std::log10(2.0));
    if (dec_digits) {
        if (fp.data_.empty()) fp.data_.push_back(0);
        fp.remove_leading_zero();
        const BigInt five_pow = fast_pow(BigInt(TEN / 2), (uint32_t)dec_digits);
        fp *= five_pow;
        const int64_t off = (int64_t)DIGIT_BITS * point_pos_ - (int64_t)dec_digits;
        BigFloat scaled(std::move(fp), -off);
        fp = BigInt(scaled, RoundMode::RoundHalfUp);
        // did rounding carry into the integer part (fp == 10^dec_digits)?
        size_type tz = 0;
        for (auto d : fp.data_) {
            if (d != 0) {
                tz += (size_type)std::countr_zero(d);
                break;
            }
            tz += DIGIT_BITS;
        }
        if (tz == dec_digits && (fp >> dec_digits) == five_pow) {
            fp.reset();
            ++ip;
        }
    }
    if (is_neg_) out.put('-');
    ip.print_dec(out);
    out.put('.');
    if (dec_digits) fp.print_dec(out, dec_digits);
    if (!direct) output << buf->rdbuf();
}

auto operator<<(std::ostream& output, const BigFloat& a) -> std::ostream& {
    a.print(output, 0, false);
    return output;
}

auto operator>>(std::istream& input, BigFloat& a) -> std::istream& {
    std::string s;
    input >> s;
    a = BigFloat(s, (input.flags() & std::ios_base::hex) != 0);
    return input;
}

// src/bigint.cpp:1044-1071
auto BigFloat::reciprocal(size_type precision) const -> BigFloat {
    if (is_zero()) throw std::domain_error("division by zero");
    if (precision == 0) precision = data_.size();
    const int64_t exponent =
        (int64_t)DIGIT_BITS * ((int64_t)data_.size() - 1 - point_pos_) + std::bit_width(data_.back());
    BigFloat b = *this >> exponent;
    b.remove_tail_zero();
    BigFloat x(1 / b.to_double());
    size_type cur = DOUBLE_MANTISSA_LEN / DIGIT_BITS;
    const BigFloat one(1);
    while (cur < precision) {
        cur *= 2;
        x += mul(x, one - mul(b, x, cur + 1), cur + 1);
    }
    x >>= exponent;
    return x;
}

// src/bigint.cpp:1074-1113: drop limbs [0, idx] rounding by mode
static void round_limbs(RoundMode mode, Digits& v, bool is_neg, int64_t idx) {
    bool up = false;
    switch (mode) {
        case RoundMode::Floor:
        case RoundMode::Ceil: {
            const bool away = (mode == RoundMode::Floor) ? is_neg : !is_neg;
            if (away) {
                int64_t i = idx;
                while (i >= 0 && v[(size_t)i] == 0) --i;
                up = i >= 0;
            }
            break;
        }
        case RoundMode::RoundHalfUp: up = (v[(size_t)idx] >> (DIGIT_BITS - 1)) != 0; break;
        case RoundMode::Truncate:
        default: up = false; break;
    }
    v.erase(v.begin(), v.begin() + idx + 1);
    if (!up) return;
    size_type i = 0;
    while (i < v.size() && v[i] == DIGIT_MASK) v[i++] = 0;
    if (i == v.size()) v.push_back(1);
    else ++v[i];
}

// src/bigint.cpp:1115-1140
void BigFloat::round(RoundMode mode, int64_t precision, RoundRelativeTo relative) {
    if (is_zero()) return;
    const int64_t n = (int64_t)data_.size();
    if (relative == RoundRelativeTo::Point) precision = n - point_pos_ - precision;
    const int64_t idx = n - precision - 1;
    if (idx >= n) {
        reset();
        return;
    }
    if (idx < 0) return;
    point_pos_ -= idx + 1;
    round_limbs(mode, data_, is_neg_, idx);
    remove_leading_zero();
}

// src/bigint.cpp:1142-1172
BigInt::BigInt(const BigFloat& x, RoundMode mode)
    : BigInt(x.is_zero() || x.point_pos_ > (int64_t)x.data_.size()
                 ? BigInt(0u)
                 : convert_from_BigFloat(mode, x.data_, x.is_neg_, x.point_pos_)) {}

BigInt::BigInt(BigFloat&& x, RoundMode mode)
    : BigInt(x.is_zero() || x.point_pos_ > (int64_t)x.data_.size()
                 ? BigInt(0u)
                 : convert_from_BigFloat(mode, std::move(x.data_), x.is_neg_, x.point_pos_)) {
    x.reset();
}

auto BigInt::convert_from_BigFloat(RoundMode mode, Digits data, bool is_neg, int64_t point_pos) -> BigInt {
    if (point_pos < 0) {
        data.insert(data.begin(), (size_type)(-point_pos), 0u);
    } else if (point_pos > 0) {
        if (point_pos > (int64_t)data.size()) return BigInt(0u);
        round_limbs(mode, data, is_neg, point_pos - 1);
    }
    BigInt r;
    r.data_ = std::move(data);
    if (r.data_.empty()) r.data_.push_back(0);
    r.is_neg_ = is_neg;
    r.remove_leading_zero();
    return r;
}

}  // namespace bigint
