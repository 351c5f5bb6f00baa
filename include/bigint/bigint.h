// BigInt / BigFloat: the host-side public API of the library, source-compatible with the
// reference's include/bigint/bigint.h (same class names, member names, signatures and
// semantics).  Every big multiply goes through bigint::mul::mul_digits (include/bigint/mul.h),
// which hands products above 16,384 bits to the MI355X engine.
//
// Semantics kept from the reference:
//   - magnitude in 28-bit limbs, canonical (no leading zero limb, zero is {0}, no -0);
//   - BigFloat value = data * 2^(-28 * point_pos), no trailing zero limbs;
//   - BigFloat::mul truncates operands and product to `precision` limbs (mul.h:223-318);
//   - reciprocal is Newton iteration with precision doubling from one limb
//     (src/bigint.cpp:1044-1071), divmod is reciprocal-based with a correction step
//     (src/bigint.cpp:625-677);
//   - init_thread_pool: the first call wins, 0 or 1 disables worker threads.
#pragma once

#include <array>
#include <compare>
#include <concepts>
#include <cstdint>
#include <iosfwd>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "bigint/bigint_base.h"
#include "bigint/mul.h"

namespace bigint {

class BigInt;
class BigFloat;

enum class RoundMode : uint8_t {
    Truncate,            // toward zero
    Floor,               // toward -infinity
    Ceil,                // toward +infinity
    RoundHalfUp,         // to nearest, ties away from zero
    Round = RoundHalfUp
};

enum class RoundRelativeTo : uint8_t {
    Significant,  // keep `precision` limbs below the most significant one
    Point         // keep `precision` limbs after the point
};

// Host worker threads (0 or 1 disables them).  The first call wins; later calls have no
// effect.  If never called, hardware_concurrency() is used.  Thread-safe.
void init_thread_pool(uint32_t n);
// Thread count fixed by the first init_thread_pool call (or the default).
auto thread_pool_size() -> uint32_t;

class BigInt {
public:
    // thresholds for divide-and-conquer decimal conversion: {from string, to string}
    static std::array<size_type, 2> DEC_STRING_BRUTE_THRESHOLDS;
    static constexpr std::array<size_type, 2> DEC_STRING_BRUTE_THRESHOLDS_DEFAULT = {5000, 2000};

    BigInt(std::unsigned_integral auto x) : BigInt(x, false) {}
    BigInt(std::signed_integral auto x) : BigInt(to_unsigned_abs(x), x < 0) {}

    BigInt(const BigInt&) = default;
    auto operator=(const BigInt&) -> BigInt& = default;
    ~BigInt() = default;
    // a moved-from object is reset to 0
    BigInt(BigInt&& o) noexcept : data_(std::move(o.data_)), is_neg_(o.is_neg_) { o.reset(); }
    auto operator=(BigInt&& o) noexcept -> BigInt& {
        if (this != &o) {
            data_ = std::move(o.data_);
            is_neg_ = o.is_neg_;
            o.reset();
        }
        return *this;
    }

    // Characters other than digits are ignored; a '-' before the first digit makes it negative.
    BigInt(const std::string& s, bool hex = false);
    explicit BigInt(const BigFloat& x, RoundMode mode = RoundMode::Truncate);
    explicit BigInt(BigFloat&& x, RoundMode mode = RoundMode::Truncate);

    [[nodiscard]] auto len() const noexcept -> size_type { return data_.size(); }
    [[nodiscard]] auto get_data() const noexcept -> const Digits& { return data_; }
    [[nodiscard]] auto is_zero() const noexcept -> bool {
        return data_.size() == 1 && data_[0] == 0;
    }
    [[nodiscard]] auto sign() const noexcept -> int {
        return is_zero() ? 0 : (is_neg_ ? -1 : 1);
    }

    [[nodiscard]] auto to_string(bool hex = false) const -> std::string {
        std::ostringstream s;
        print(s, hex, true);
        return s.str();
    }
    // hex prints no "0x" prefix but does print '-'
    void print(std::ostream& output, bool hex = false, bool direct = false) const;

    void reset() noexcept {
        data_.resize(1);
        data_[0] = 0;
        is_neg_ = false;
    }
    void flip_sign() noexcept {
        if (!is_zero()) is_neg_ = !is_neg_;
    }
    void remove_sign() noexcept { is_neg_ = false; }

    static auto compare_abs(const BigInt& a, const BigInt& b) noexcept -> std::strong_ordering;

    // |this| /= b in place, returns |this| mod b
    auto unsigned_inplace_divmod(uint64_t b) -> uint64_t;

    auto operator+() const -> BigInt { return *this; }
    auto operator-() const -> BigInt {
        BigInt r = *this;
        r.flip_sign();
        return r;
    }

    auto operator++() -> BigInt&;
    [[nodiscard]] auto operator++(int) -> BigInt;
    auto operator--() -> BigInt&;
    [[nodiscard]] auto operator--(int) -> BigInt;

    auto operator+=(const BigInt& b) -> BigInt& {
        inplace_add_or_sub(b, false);
        return *this;
    }
    auto operator+=(Integer64 auto b) -> BigInt& {
        inplace_add_or_sub((uint64_t)to_unsigned_abs(b), b < 0);
        return *this;
    }
    auto operator-=(const BigInt& b) -> BigInt& {
        inplace_add_or_sub(b, true);
        return *this;
    }
    auto operator-=(Integer64 auto b) -> BigInt& {
        inplace_add_or_sub((uint64_t)to_unsigned_abs(b), b > 0);
        return *this;
    }

    auto operator*=(const BigInt& b) -> BigInt& { return *this = *this * b; }
    auto operator*=(Integer64 auto b) -> BigInt& {
        unsigned_inplace_mul((uint64_t)to_unsigned_abs(b));
        if (b < 0) flip_sign();
        return *this;
    }
    auto operator/=(Integer64 auto b) -> BigInt& {
        if (b == 0) throw std::domain_error("division by zero");
        if (is_zero()) return *this;
        unsigned_inplace_divmod((uint64_t)to_unsigned_abs(b));
        if (b < 0) flip_sign();
        return *this;
    }

    // negative offsets shift the other way
    auto operator<<=(Integer64 auto offset) -> BigInt& {
        if (offset > 0) shift_left(data_, (uint64_t)offset);
        else if (offset < 0) shift_right(data_, (uint64_t)to_unsigned_abs(offset));
        remove_leading_zero();
        return *this;
    }
    auto operator>>=(Integer64 auto offset) -> BigInt& {
        if (offset > 0) shift_right(data_, (uint64_t)offset);
        else if (offset < 0) shift_left(data_, (uint64_t)to_unsigned_abs(offset));
        remove_leading_zero();
        return *this;
    }

    // bitwise operations ignore the sign
    auto operator&=(const BigInt& b) -> BigInt&;
    auto operator|=(const BigInt& b) -> BigInt&;
    auto operator^=(const BigInt& b) -> BigInt&;
    // complement over `len` limbs (0 = own length)
    auto bitwise_not(size_type len = 0) -> BigInt&;

    explicit operator bool() const noexcept { return !is_zero(); }

    friend auto operator<=>(const BigInt& a, Literal_zero) noexcept -> std::strong_ordering {
        return a.sign() <=> 0;
    }
    friend auto operator<=>(Literal_zero, const BigInt& b) noexcept -> std::strong_ordering {
        return 0 <=> b.sign();
    }
    friend auto operator==(const BigInt& a, Literal_zero) noexcept -> bool { return a.is_zero(); }
    friend auto operator==(Literal_zero, const BigInt& b) noexcept -> bool { return b.is_zero(); }
    friend auto operator<=>(const BigInt& a, const BigInt& b) noexcept -> std::strong_ordering {
        const int sa = a.sign(), sb = b.sign();
        if (auto c = sa <=> sb; c != 0) return c;
        if (sa == 0) return std::strong_ordering::equal;
        const auto c = compare_abs(a, b);
        return sa < 0 ? 0 <=> c : c;
    }
    friend auto operator==(const BigInt& a, const BigInt& b) noexcept -> bool {
        return a.sign() == b.sign() && compare_abs(a, b) == 0;
    }

    friend auto operator<<(std::ostream& output, const BigInt& a) -> std::ostream&;
    friend auto operator>>(std::istream& input, BigInt& a) -> std::istream&;

    friend auto operator+(BigInt a, const BigInt& b) -> BigInt { return a += b; }
    friend auto operator+(BigInt a, Integer64 auto b) -> BigInt { return a += b; }
    friend auto operator+(Integer64 auto a, BigInt b) -> BigInt { return b += a; }
    friend auto operator-(BigInt a, const BigInt& b) -> BigInt { return a -= b; }
    friend auto operator-(BigInt a, Integer64 auto b) -> BigInt { return a -= b; }
    friend auto operator-(Integer64 auto a, BigInt b) -> BigInt {
        b -= a;
        b.flip_sign();
        return b;
    }

    // The big product: zero short-circuit, sign by xor, magnitude from mul_digits (which
    // squares when a and b are the same object).
    friend auto operator*(const BigInt& a, const BigInt& b) -> BigInt {
        if (a.is_zero() || b.is_zero()) return 0;
        BigInt r;
        r.data_ = mul::mul_digits(a.data_, b.data_);
        r.is_neg_ = a.is_neg_ != b.is_neg_;
        r.remove_leading_zero();
        return r;
    }
    friend auto operator*(BigInt a, Integer64 auto b) -> BigInt { return a *= b; }
    friend auto operator*(Integer64 auto a, BigInt b) -> BigInt { return b *= a; }
    friend auto operator/(BigInt a, Integer64 auto b) -> BigInt { return a /= b; }
    friend auto operator<<(BigInt a, Integer64 auto offset) -> BigInt { return a <<= offset; }
    friend auto operator>>(BigInt a, Integer64 auto offset) -> BigInt { return a >>= offset; }
    friend auto operator&(BigInt a, const BigInt& b) -> BigInt { return a &= b; }
    friend auto operator|(BigInt a, const BigInt& b) -> BigInt { return a |= b; }
    friend auto operator^(BigInt a, const BigInt& b) -> BigInt { return a ^= b; }

    // quotient rounded by `mode`, and remainder a - q*b
    [[nodiscard]] auto divmod(const BigInt& b, RoundMode mode = RoundMode::Truncate,
                              bool check = true) const -> std::pair<BigInt, BigInt>;

    static auto get_pow_of_ten(uint32_t exponent) -> BigInt;

    // The integer with magnitude sum digits[i] 2^(28 i) (every limb < 2^28) and the given sign,
    // canonicalized (no leading zero limbs, no negative zero).  An extension of the reference
    // API for callers that build numbers limb by limb.
    static auto from_digits(Digits digits, bool negative = false) -> BigInt;

    // a_i * b_i for every pair, each equal to operator*; the big ones run as one grouped GPU
    // launch set (one product-tree level at a time).
    static auto mul_batch(const std::vector<std::pair<const BigInt*, const BigInt*>>& pairs,
                          int plan_mode = 1) -> std::vector<BigInt>;

    friend class BigFloat;

private:
    Digits data_;
    bool is_neg_;

    explicit BigInt() : is_neg_(false) {}
    explicit BigInt(std::unsigned_integral auto value, bool is_neg) : is_neg_(is_neg) {
        if (value == 0) data_.push_back(0);
        while (value) {
            data_.push_back(static_cast<uint32_t>(value) & DIGIT_MASK);
            value >>= DIGIT_BITS;
        }
        if (data_.size() == 1 && data_[0] == 0) is_neg_ = false;
    }

    static auto convert_from_dec_string(const std::string& s, size_type start, size_type end)
        -> std::pair<BigInt, size_type>;
    [[nodiscard]] auto to_dec_string_brute() const -> std::string;
    void print_dec(std::ostream& output, size_type len = 0) const;
    void print_hex(std::ostream& output) const;
    auto remove_leading_zero() -> size_type;
    void unsigned_self_inc_or_dec(bool is_inc);
    void inplace_add_or_sub(const BigInt& b, bool is_sub);
    void inplace_add_or_sub(uint64_t b, bool is_sub);
    void unsigned_inplace_mul(uint64_t b);
    static void shift_left(Digits& v, uint64_t offset);
    static void shift_right(Digits& v, uint64_t offset);
    static auto convert_from_BigFloat(RoundMode mode, Digits data, bool is_neg, int64_t point_pos)
        -> BigInt;
};

class BigFloat {
public:
    explicit BigFloat(const BigInt& x, int64_t offset = 0);
    explicit BigFloat(BigInt&& x, int64_t offset = 0);
    BigFloat(std::integral auto value, int64_t offset = 0) : BigFloat(BigInt(value), offset) {}
    // the string is an integer (no decimal point support)
    BigFloat(const std::string& s, bool hex = false, int64_t offset = 0)
        : BigFloat(BigInt(s, hex), offset) {}
    BigFloat(double value);

    BigFloat(const BigFloat&) = default;
    auto operator=(const BigFloat&) -> BigFloat& = default;
    ~BigFloat() = default;
    BigFloat(BigFloat&& o) noexcept
        : data_(std::move(o.data_)), point_pos_(o.point_pos_), is_neg_(o.is_neg_) {
        o.reset();
    }
    auto operator=(BigFloat&& o) noexcept -> BigFloat& {
        if (this != &o) {
            data_ = std::move(o.data_);
            point_pos_ = o.point_pos_;
            is_neg_ = o.is_neg_;
            o.reset();
        }
        return *this;
    }

    [[nodiscard]] auto len() const noexcept -> size_type { return data_.size(); }
    [[nodiscard]] auto get_point_pos() const noexcept -> int64_t { return point_pos_; }
    [[nodiscard]] auto get_data() const noexcept -> const Digits& { return data_; }
    [[nodiscard]] auto is_zero() const noexcept -> bool {
        return data_.size() == 1 && data_[0] == 0;
    }
    [[nodiscard]] auto sign() const noexcept -> int {
        return is_zero() ? 0 : (is_neg_ ? -1 : 1);
    }
    [[nodiscard]] auto to_double() const -> double;

    void reset() noexcept {
        data_.resize(1);
        data_[0] = 0;
        point_pos_ = 0;
        is_neg_ = false;
    }
    void flip_sign() noexcept {
        if (!is_zero()) is_neg_ = !is_neg_;
    }
    void remove_sign() noexcept { is_neg_ = false; }

    // decimal text with `dec_digits` fraction digits (0 = as many as the binary fraction holds)
    void print(std::ostream& output, size_type dec_digits = 0, bool direct = false) const;
    friend auto operator<<(std::ostream& output, const BigFloat& a) -> std::ostream&;
    [[nodiscard]] auto to_string(size_type dec_digits = 0) const -> std::string {
        std::ostringstream s;
        print(s, dec_digits, true);
        return s.str();
    }
    friend auto operator>>(std::istream& input, BigFloat& a) -> std::istream&;

    // 1 / this to `precision` limbs (0 = own length)
    [[nodiscard]] auto reciprocal(size_type precision = 0) const -> BigFloat;

    [[nodiscard]] auto operator+() const -> BigFloat { return *this; }
    [[nodiscard]] auto operator-() const -> BigFloat {
        BigFloat r = *this;
        r.flip_sign();
        return r;
    }

    void round(RoundMode mode, int64_t precision, RoundRelativeTo relative);

    auto operator<<=(int64_t offset) -> BigFloat& {
        shift(offset);
        return *this;
    }
    auto operator>>=(int64_t offset) -> BigFloat& {
        shift(-offset);
        return *this;
    }
    auto operator+=(const BigFloat& b) -> BigFloat& { return *this = *this + b; }
    auto operator-=(const BigFloat& b) -> BigFloat& { return *this = *this - b; }
    auto operator*=(Integer64 auto b) -> BigFloat& {
        unsigned_inplace_mul((uint64_t)to_unsigned_abs(b));
        if (b < 0) flip_sign();
        return *this;
    }
    auto operator*=(const BigFloat& b) -> BigFloat& { return *this = *this * b; }

    friend auto operator<<(BigFloat x, int64_t offset) -> BigFloat { return x <<= offset; }
    friend auto operator>>(BigFloat x, int64_t offset) -> BigFloat { return x >>= offset; }
    friend auto operator+(const BigFloat& a, const BigFloat& b) -> BigFloat {
        return add_or_sub(a, b, false);
    }
    friend auto operator-(const BigFloat& a, const BigFloat& b) -> BigFloat {
        return add_or_sub(a, b, true);
    }
    friend auto operator*(const BigFloat& a, const BigFloat& b) -> BigFloat {
        if (a.is_zero() || b.is_zero()) return 0;
        return mul(a, b, 0);
    }
    friend auto operator*(BigFloat a, Integer64 auto b) -> BigFloat { return a *= b; }
    friend auto operator*(Integer64 auto a, BigFloat b) -> BigFloat { return b *= a; }

    // product truncated to `precision` limbs (0 = exact)
    static auto mul(const BigFloat& a, const BigFloat& b, size_type precision) -> BigFloat {
        if (a.is_zero() || b.is_zero()) return 0;
        BigFloat r;
        r.point_pos_ = a.point_pos_ + b.point_pos_;
        r.data_ = mul::mul_digits(a.data_, b.data_, precision, &r.point_pos_);
        r.is_neg_ = a.is_neg_ != b.is_neg_;
        r.remove_leading_zero();
        r.remove_tail_zero();
        return r;
    }

    friend class BigInt;

private:
    Digits data_;
    int64_t point_pos_;
    bool is_neg_;

    explicit BigFloat() : point_pos_(0), is_neg_(false) {}
    void shift(int64_t offset);
    static void shift_right(Digits& v, uint32_t offset);
    static void shift_left(Digits& v, uint32_t offset);
    auto remove_leading_zero() -> size_type;
    auto remove_tail_zero() -> size_type;
    static auto add_or_sub(const BigFloat& a, const BigFloat& b, bool is_sub) -> BigFloat;
    void unsigned_inplace_mul(uint64_t b);
};

template <class T>
concept can_remove_sign = requires(T x) {
    { x.remove_sign() };
};

template <can_remove_sign T>
auto abs(const T& x) -> T {
    T r = x;
    r.remove_sign();
    return r;
}

struct BigIntPrintHelper {
    const BigInt* ptr;
    bool hex;
    bool direct;
    friend auto operator<<(std::ostream& output, const BigIntPrintHelper& h) -> std::ostream& {
        h.ptr->print(output, h.hex, h.direct);
        return output;
    }
};

struct BigFloatPrintHelper {
    const BigFloat* ptr;
    size_type dec_digits;
    bool direct;
    friend auto operator<<(std::ostream& output, const BigFloatPrintHelper& h) -> std::ostream& {
        h.ptr->print(output, h.dec_digits, h.direct);
        return output;
    }
};

inline auto print(const BigInt& x, bool hex = false, bool direct = false) -> BigIntPrintHelper {
    return {.ptr = &x, .hex = hex, .direct = direct};
}
inline auto print(const BigFloat& x, size_type dec_digits = 0, bool direct = false)
    -> BigFloatPrintHelper {
    return {.ptr = &x, .dec_digits = dec_digits, .direct = direct};
}

}  // namespace bigint
