// Host-side entry points of the plan-selection MLP kernels (planner_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bi {

constexpr int kMaxPlans = 8;

int mlp_param_count(int K);
hipError_t mlp_forward(const float* x, int B, const float* p, int K, float* y, float* h1,
                       float* h2, hipStream_t st);
// Forward + backward of the masked MSE.  grad holds NP + 1 floats: the gradient, then the loss.
// y, h1, h2 are B*K, B*64, B*64 scratch; part / part_loss are ceil(B/64)*NP and ceil(B/64).
hipError_t mlp_grad(const float* x, const float* t, const float* mask, int B, const float* p,
                    int K, float inv_count, float* y, float* h1, float* h2, float* part,
                    float* part_loss, float* grad, hipStream_t st);
hipError_t mlp_adam(float* p, const float* g, float* m, float* v, int NP, float lr, float b1,
                    float b2, float eps, int step, hipStream_t st);
// Per sample the allowed plan minimising y + offset (log2 of the predicted time per product),
// unless prior[s] (the table's or the heuristic's plan, -1 for none) is allowed and either scores
// within `margin` of it or, with log2_batch (log2 of each sample's batch size), is predicted to
// take under 2^min_log2_us for the whole batch: then prior[s].
hipError_t mlp_choose(const float* y, const float* offset, const uint8_t* allowed, const int32_t* prior,
                      float margin, const float* log2_batch, float min_log2_us, int B, int K,
                      int32_t* choice, hipStream_t st);

}  // namespace bi
