// Signed integers kept in the multiply backend's memory (HBM on the MI355X engine) between
// batched operations, for product trees whose levels never return to the host.  An extension of
// the reference API: the reference keeps every number in host memory and merges a tree level
// with BigInt operator* and operator+ (examples/pi.cpp:13-80).
#pragma once

#include <memory>
#include <utility>
#include <vector>

#include "bigint/bigint.h"

namespace bigint {

class DeviceInt {
  public:
    DeviceInt() = default;

    // Many numbers to the backend in one transfer, and back.
    static auto upload(const std::vector<const BigInt*>& xs) -> std::vector<DeviceInt>;
    static auto download(const std::vector<const DeviceInt*>& xs) -> std::vector<BigInt>;

    // a_i * b_i for every pair as one grouped batch (the same pointer twice squares), equal to
    // BigInt operator* of the values.
    static auto mul_batch(const std::vector<std::pair<const DeviceInt*, const DeviceInt*>>& pairs,
                          int plan_mode = 1) -> std::vector<DeviceInt>;
    // a_i + b_i (signed) for every pair as one batch, equal to BigInt operator+.
    static auto add_batch(const std::vector<std::pair<const DeviceInt*, const DeviceInt*>>& pairs)
        -> std::vector<DeviceInt>;

    // limbs in the backend, at least the value's length (a product may carry one zero top limb)
    [[nodiscard]] auto len() const noexcept -> size_type { return len_; }
    [[nodiscard]] auto is_neg() const noexcept -> bool { return neg_; }
    [[nodiscard]] auto is_zero() const noexcept -> bool { return len_ == 0; }

  private:
    std::shared_ptr<void> buf_;  // the backend allocation holding the limbs (shared by a batch)
    const uint32_t* limbs_ = nullptr;
    size_type len_ = 0;          // 0 for zero
    bool neg_ = false;
};

}  // namespace bigint
