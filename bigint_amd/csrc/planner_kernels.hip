// Learned plan selection: a small MLP cost model trained and evaluated on the GPU.
//
//   x = (log2 la, log2 lb, log2 batch), normalised      (3 inputs)
//   h1 = relu(x W1 + b1)      64
//   h2 = relu(h1 W2 + b2)     64
//   y  = h2 W3 + b3           K  (predicted log2 time per product, one output per plan)
//
// Parameters are one fp32 vector laid out as W1[3][64] b1[64] W2[64][64] b2[64] W3[64][K] b3[K].
// Training minimises the masked mean squared error over measured (sample, plan) entries.  The
// gradient is reduced deterministically (per-block partial sums, then a fixed-order reduction) so
// it can be compared with a float64 CPU reference.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "planner.h"

namespace bi {

constexpr int kIn = 3;
constexpr int kHid = 64;
constexpr int kGradSamples = 64;  // samples per gradient block

__host__ __device__ constexpr int mlp_nparams(int K) {
    return kIn * kHid + kHid + kHid * kHid + kHid + kHid * K + K;
}

struct MlpView {
    const float *W1, *b1, *W2, *b2, *W3, *b3;
};
__device__ __forceinline__ MlpView view(const float* p, int K) {
    MlpView v;
    v.W1 = p;
    v.b1 = v.W1 + kIn * kHid;
    v.W2 = v.b1 + kHid;
    v.b2 = v.W2 + kHid * kHid;
    v.W3 = v.b2 + kHid;
    v.b3 = v.W3 + kHid * K;
    return v;
}

// One block of kHid threads per sample: thread j computes hidden unit j of both layers and, for
// j < K, output j, each as a sequential fmaf chain over its inputs in index order; the layer
// activations are shared through LDS.  Optionally keeps the activations.
__global__ __launch_bounds__(kHid) void k_mlp_forward(const float* __restrict__ x, int B,
                                                      const float* __restrict__ p, int K,
                                                      float* __restrict__ y, float* __restrict__ h1o,
                                                      float* __restrict__ h2o) {
    __shared__ float sh1[kHid], sh2[kHid];
    const int s = blockIdx.x;
    const int j = threadIdx.x;
    if (s >= B) return;
    const MlpView v = view(p, K);
    float a = v.b1[j];
#pragma unroll
    for (int i = 0; i < kIn; ++i) a = fmaf(x[s * kIn + i], v.W1[i * kHid + j], a);
    const float h1 = a > 0.f ? a : 0.f;
    sh1[j] = h1;
    __syncthreads();
    float c = v.b2[j];
#pragma unroll 16
    for (int i = 0; i < kHid; ++i) c = fmaf(sh1[i], v.W2[i * kHid + j], c);
    const float h2 = c > 0.f ? c : 0.f;
    sh2[j] = h2;
    __syncthreads();
    if (j < K) {
        float o = v.b3[j];
#pragma unroll 16
        for (int i = 0; i < kHid; ++i) o = fmaf(sh2[i], v.W3[i * K + j], o);
        y[s * K + j] = o;
    }
    if (h1o) {
        h1o[(size_t)s * kHid + j] = h1;
        h2o[(size_t)s * kHid + j] = h2;
    }
}

// Per-block partial gradients over kGradSamples samples.  dy = 2 * mask * (y - t) / count.
// Output: part[block][nparams], plus part_loss[block].
__global__ __launch_bounds__(256) void k_mlp_grad_partial(
    const float* __restrict__ x, const float* __restrict__ t, const float* __restrict__ mask,
    int B, const float* __restrict__ p, int K, const float* __restrict__ y,
    const float* __restrict__ h1, const float* __restrict__ h2, float inv_count,
    float* __restrict__ part, float* __restrict__ part_loss) {
    __shared__ float s_h1[kGradSamples][kHid];
    __shared__ float s_dh1[kGradSamples][kHid];
    __shared__ float s_h2[kGradSamples][kHid];
    __shared__ float s_dh2[kGradSamples][kHid];
    __shared__ float s_dy[kGradSamples][8];
    __shared__ float s_x[kGradSamples][kIn];
    __shared__ float s_loss[kGradSamples];
    const MlpView v = view(p, K);
    const int NP = mlp_nparams(K);
    const int s0 = blockIdx.x * kGradSamples;
    const int tid = threadIdx.x;
    // stage activations and output gradients
    for (int e = tid; e < kGradSamples * kHid; e += blockDim.x) {
        const int ls = e / kHid, j = e % kHid, s = s0 + ls;
        s_h1[ls][j] = s < B ? h1[(size_t)s * kHid + j] : 0.f;
        s_h2[ls][j] = s < B ? h2[(size_t)s * kHid + j] : 0.f;
    }
    for (int e = tid; e < kGradSamples * 8; e += blockDim.x) {
        const int ls = e / 8, k = e % 8, s = s0 + ls;
        float d = 0.f;
        if (s < B && k < K) {
            const float m = mask[s * K + k];
            d = m * 2.f * (y[s * K + k] - t[s * K + k]) * inv_count;
        }
        s_dy[ls][k] = d;
    }
    for (int e = tid; e < kGradSamples * kIn; e += blockDim.x) {
        const int ls = e / kIn, i = e % kIn, s = s0 + ls;
        s_x[ls][i] = s < B ? x[s * kIn + i] : 0.f;
    }
    if (tid < kGradSamples) {
        const int s = s0 + tid;
        float l = 0.f;
        if (s < B)
            for (int k = 0; k < K; ++k) {
                const float d = y[s * K + k] - t[s * K + k];
                l += mask[s * K + k] * d * d;
            }
        s_loss[tid] = l * inv_count;
    }
    __syncthreads();
    // dh2 = (dy W3^T) * (h2 > 0)
    for (int e = tid; e < kGradSamples * kHid; e += blockDim.x) {
        const int ls = e / kHid, j = e % kHid;
        float a = 0.f;
        for (int k = 0; k < K; ++k) a = fmaf(s_dy[ls][k], v.W3[j * K + k], a);
        s_dh2[ls][j] = s_h2[ls][j] > 0.f ? a : 0.f;
    }
    __syncthreads();
    // dh1 = (dh2 W2^T) * (h1 > 0)
    for (int e = tid; e < kGradSamples * kHid; e += blockDim.x) {
        const int ls = e / kHid, i = e % kHid;
        float a = 0.f;
        for (int j = 0; j < kHid; ++j) a = fmaf(s_dh2[ls][j], v.W2[i * kHid + j], a);
        s_dh1[ls][i] = s_h1[ls][i] > 0.f ? a : 0.f;
    }
    __syncthreads();
    float* out = part + (size_t)blockIdx.x * NP;
    float* gW1 = out;
    float* gb1 = gW1 + kIn * kHid;
    float* gW2 = gb1 + kHid;
    float* gb2 = gW2 + kHid * kHid;
    float* gW3 = gb2 + kHid;
    float* gb3 = gW3 + kHid * K;
    for (int e = tid; e < kIn * kHid; e += blockDim.x) {
        const int i = e / kHid, j = e % kHid;
        float a = 0.f;
        for (int ls = 0; ls < kGradSamples; ++ls) a = fmaf(s_x[ls][i], s_dh1[ls][j], a);
        gW1[e] = a;
    }
    for (int j = tid; j < kHid; j += blockDim.x) {
        float a1 = 0.f, a2 = 0.f;
        for (int ls = 0; ls < kGradSamples; ++ls) {
            a1 += s_dh1[ls][j];
            a2 += s_dh2[ls][j];
        }
        gb1[j] = a1;
        gb2[j] = a2;
    }
    for (int e = tid; e < kHid * kHid; e += blockDim.x) {
        const int i = e / kHid, j = e % kHid;
        float a = 0.f;
        for (int ls = 0; ls < kGradSamples; ++ls) a = fmaf(s_h1[ls][i], s_dh2[ls][j], a);
        gW2[e] = a;
    }
    for (int e = tid; e < kHid * K; e += blockDim.x) {
        const int i = e / K, k = e % K;
        float a = 0.f;
        for (int ls = 0; ls < kGradSamples; ++ls) a = fmaf(s_h2[ls][i], s_dy[ls][k], a);
        gW3[e] = a;
    }
    for (int k = tid; k < K; k += blockDim.x) {
        float a = 0.f;
        for (int ls = 0; ls < kGradSamples; ++ls) a += s_dy[ls][k];
        gb3[k] = a;
    }
    if (tid == 0) {
        float l = 0.f;
        for (int ls = 0; ls < kGradSamples; ++ls) l += s_loss[ls];
        part_loss[blockIdx.x] = l;
    }
}

// Fixed-order sum of the per-block partials.
__global__ void k_reduce_partials(const float* __restrict__ part, int nblocks, int NP,
                                  float* __restrict__ grad) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NP) return;
    float a = 0.f;
    for (int b = 0; b < nblocks; ++b) a += part[(size_t)b * NP + e];
    grad[e] = a;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int NP, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NP) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Plan choice: argmin of the predicted cost over the plans allowed for each sample.
__global__ void k_argmin_plan(const float* __restrict__ y, const float* __restrict__ offset,
                              const uint8_t* __restrict__ allowed, const int32_t* __restrict__ prior,
                              float margin, const float* __restrict__ log2_batch, float min_log2_us,
                              int B, int K, int32_t* __restrict__ choice) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    int best = -1;
    float bv = 0.f;
    for (int k = 0; k < K; ++k) {
        if (!allowed[s * K + k]) continue;
        const float c = y[s * K + k] + (offset ? offset[s * K + k] : 0.f);
        if (best < 0 || c < bv) {
            best = k;
            bv = c;
        }
    }
    // the prior plan stands unless the model predicts a gain over it larger than the margin, and
    // (with log2_batch) unless its predicted time for the whole batch is below 2^min_log2_us
    const int pr = prior ? prior[s] : -1;
    if (best >= 0 && pr >= 0 && pr < K && allowed[s * K + pr]) {
        const float cp = y[s * K + pr] + (offset ? offset[s * K + pr] : 0.f);
        if (cp - bv < margin || (log2_batch && cp + log2_batch[s] < min_log2_us)) best = pr;
    }
    choice[s] = best < 0 ? 0 : best;
}

// ---- host wrappers ---------------------------------------------------------------------------
int mlp_param_count(int K) { return mlp_nparams(K); }

hipError_t mlp_forward(const float* x, int B, const float* p, int K, float* y, float* h1,
                       float* h2, hipStream_t st) {
    if (B <= 0) return hipSuccess;
    (void)hipGetLastError();  // report this launch's error, not an earlier ignored one
    static_assert(kMaxPlans <= kHid, "one output per hidden-unit thread");
    hipLaunchKernelGGL(k_mlp_forward, dim3(B), dim3(kHid), 0, st, x, B, p, K, y, h1, h2);
    return hipGetLastError();
}

hipError_t mlp_grad(const float* x, const float* t, const float* mask, int B, const float* p,
                    int K, float inv_count, float* y, float* h1, float* h2, float* part,
                    float* part_loss, float* grad, hipStream_t st) {
    hipError_t e = mlp_forward(x, B, p, K, y, h1, h2, st);
    if (e != hipSuccess) return e;
    const int nb = (B + kGradSamples - 1) / kGradSamples;
    (void)hipGetLastError();  // report this launch's error, not an earlier ignored one
    hipLaunchKernelGGL(k_mlp_grad_partial, dim3(nb), dim3(256), 0, st, x, t, mask, B, p, K, y, h1,
                       h2, inv_count, part, part_loss);
    const int NP = mlp_nparams(K);
    (void)hipGetLastError();  // report this launch's error, not an earlier ignored one
    hipLaunchKernelGGL(k_reduce_partials, dim3((NP + 255) / 256), dim3(256), 0, st, part, nb, NP, grad);
    // loss: reduce part_loss into grad[NP] (caller allocates NP + 1)
    (void)hipGetLastError();  // report this launch's error, not an earlier ignored one
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1), 0, st, part_loss, nb, 1, grad + NP);
    return hipGetLastError();
}

hipError_t mlp_adam(float* p, const float* g, float* m, float* v, int NP, float lr, float b1,
                    float b2, float eps, int step, hipStream_t st) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    (void)hipGetLastError();  // report this launch's error, not an earlier ignored one
    hipLaunchKernelGGL(k_adam, dim3((NP + 255) / 256), dim3(256), 0, st, p, g, m, v, NP, lr, b1, b2,
                       eps, c1, c2);
    return hipGetLastError();
}

hipError_t mlp_choose(const float* y, const float* offset, const uint8_t* allowed, const int32_t* prior,
                      float margin, const float* log2_batch, float min_log2_us, int B, int K,
                      int32_t* choice, hipStream_t st) {
    if (B <= 0) return hipSuccess;
    (void)hipGetLastError();  // report this launch's error, not an earlier ignored one
    hipLaunchKernelGGL(k_argmin_plan, dim3((B + 255) / 256), dim3(256), 0, st, y, offset, allowed, prior,
                       margin, log2_batch, min_log2_us, B, K, choice);
    return hipGetLastError();
}

}  // namespace bi
