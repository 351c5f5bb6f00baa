// Base types of the host library: the reference's limb layout and small integer helpers.
// Layout: Digits = std::vector<uint32_t>, little-endian, 28 bits per limb
// (reference include/bigint/bigint_base.h:21-26).
#pragma once

#include <bit>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <type_traits>
#include <vector>

namespace bigint {

using Digits = std::vector<uint32_t>;
using size_type = std::size_t;
using uint128_t = __uint128_t;

constexpr uint32_t DIGIT_BITS = 28;
constexpr uint32_t DIGIT_MASK = (1u << DIGIT_BITS) - 1u;

constexpr int TEN = 10;
constexpr int DOUBLE_MANTISSA_LEN = 52;
constexpr int DOUBLE_EXPONENT_LEN = 11;
constexpr int DOUBLE_EXPONENT_BIAS = 1023;

// compile-time literal zero for fast comparisons (`x == 0`, `x < 0`)
using Literal_zero = decltype(nullptr);

// Integral types of at most 64 bits, bool excluded.
template <typename T>
concept Integer64 = std::is_integral_v<T> && !std::is_same_v<std::remove_cv_t<T>, bool> &&
                    (std::numeric_limits<T>::digits <= 64);

template <std::integral T>
[[nodiscard]] constexpr auto to_unsigned_abs(T x) -> std::make_unsigned_t<T> {
    using U = std::make_unsigned_t<T>;
    return x >= 0 ? static_cast<U>(x) : U(0) - static_cast<U>(x);
}

// Square-and-multiply power for any type with *= (reference bigint_base.h:60-69).
template <typename T>
constexpr auto fast_pow(T base, uint32_t exponent) -> T {
    T res = 1;
    while (exponent) {
        if (exponent & 1u) res *= base;
        exponent >>= 1;
        if (exponent) base *= base;
    }
    return res;
}

template <std::unsigned_integral T>
constexpr auto ceil_div(T a, T b) -> T {
    return a / b + (a % b != 0);
}

}  // namespace bigint
