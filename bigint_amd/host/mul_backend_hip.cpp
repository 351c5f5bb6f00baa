// Product backend of the host library: the MI355X engine through the C ABI (include/bigint_hip.h).
#include <new>

#include "bigint/mul.h"
#include "bigint_hip.h"
#include "mul_backend.h"

namespace bigint::mul::backend {

void multiply(std::vector<Job>& jobs, int plan_mode) {
    if (jobs.size() == 1) {
        Job& j = jobs[0];
        size_t len = 0;
        const int rc = bi_hip_mul(j.a.data(), j.a.size(), j.b.data(), j.b.size(), j.square ? 1 : 0,
                                  j.out, j.a.size() + j.b.size(), &len);
        if (rc == BI_HIP_ENOMEM) throw std::bad_alloc();
        if (rc != BI_HIP_OK) throw engine_error(rc, bi_hip_strerror(rc));
        j.out_len = len;
        return;
    }
    std::vector<bi_hip_pair> hp(jobs.size());
    for (size_t k = 0; k < jobs.size(); ++k) {
        const Job& j = jobs[k];
        hp[k] = bi_hip_pair{j.a.data(), j.a.size(), j.square ? j.a.data() : j.b.data(),
                            j.square ? j.a.size() : j.b.size(), j.out, j.a.size() + j.b.size(), 0, 0};
    }
    const int rc = bi_hip_mul_batch(hp.data(), hp.size(), plan_mode);
    if (rc == BI_HIP_ENOMEM) throw std::bad_alloc();
    if (rc != BI_HIP_OK) throw engine_error(rc, bi_hip_strerror(rc));
    for (size_t k = 0; k < jobs.size(); ++k) jobs[k].out_len = hp[k].out_len;
}

auto multiply_top(std::span<const uint32_t> a, std::span<const uint32_t> b, bool square, size_t keep,
                  uint32_t* out) -> size_t {
    size_t len = 0;
    const int rc = bi_hip_mul_top(a.data(), a.size(), square ? a.data() : b.data(),
                                  square ? a.size() : b.size(), square ? 1 : 0, keep, out, &len);
    if (rc == BI_HIP_ENOMEM) throw std::bad_alloc();
    if (rc != BI_HIP_OK) throw engine_error(rc, bi_hip_strerror(rc));
    return len;
}

bool load_planner(const char* path) { return bi_hip_load_planner(path) == BI_HIP_OK; }

}  // namespace bigint::mul::backend
