// Product backend of the host library: the MI355X engine through the C ABI (include/bigint_hip.h).
#include <algorithm>
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

namespace {
void check(int rc) {
    if (rc == BI_HIP_ENOMEM) throw std::bad_alloc();
    if (rc != BI_HIP_OK) throw engine_error(rc, bi_hip_strerror(rc));
}
}  // namespace

auto dev_alloc(size_t words) -> std::shared_ptr<void> {
    void* p = nullptr;
    check(bi_hip_device_alloc(words * 4, &p));
    // product batches run asynchronously on the engine's stream: wait for them before a buffer
    // they may still read or write goes back
    return std::shared_ptr<void>(p, [](void* q) {
        (void)bi_hip_synchronize();
        (void)bi_hip_device_free(q);
    });
}

void dev_upload(uint32_t* dst, const std::vector<std::span<const uint32_t>>& src) {
    std::vector<const void*> p(src.size());
    std::vector<size_t> n(src.size());
    for (size_t i = 0; i < src.size(); ++i) {
        p[i] = src[i].data();
        n[i] = src[i].size() * 4;
    }
    check(bi_hip_copy_to_device(dst, p.data(), n.data(), p.size()));
}

void dev_download(const std::vector<std::span<uint32_t>>& out, const uint32_t* src) {
    std::vector<void*> p(out.size());
    std::vector<size_t> n(out.size());
    for (size_t i = 0; i < out.size(); ++i) {
        p[i] = out[i].data();
        n[i] = out[i].size() * 4;
    }
    check(bi_hip_copy_to_host(p.data(), n.data(), p.size(), src));
}

void dev_mul_batch(const std::vector<DevMul>& jobs, int plan_mode) {
    std::vector<bi_hip_dpair> dp(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i) {
        const DevMul& j = jobs[i];
        dp[i] = bi_hip_dpair{j.a, j.la, j.square ? j.a : j.b, j.square ? j.la : j.lb, j.out, j.la + j.lb};
    }
    check(bi_hip_mul_batch_device(dp.data(), dp.size(), plan_mode, nullptr));
}

void dev_add_batch(std::vector<DevAdd>& jobs) {
    constexpr size_t kMaxJobs = 65535;  // one grid row per job
    for (size_t lo = 0; lo < jobs.size(); lo += kMaxJobs) {
        const size_t hi = std::min(jobs.size(), lo + kMaxJobs);
        std::vector<bi_hip_dadd> d(hi - lo);
        for (size_t i = lo; i < hi; ++i) {
            const DevAdd& j = jobs[i];
            d[i - lo] = bi_hip_dadd{j.x, j.lx, j.y, j.ly, j.subtract ? 1 : 0, j.out,
                                    std::max(j.lx, j.ly) + 1, 0, 0};
        }
        check(bi_hip_add_batch_device(d.data(), d.size(), nullptr));
        for (size_t i = lo; i < hi; ++i) {
            jobs[i].swapped = d[i - lo].swapped != 0;
            jobs[i].out_len = d[i - lo].out_len;
        }
    }
}

}  // namespace bigint::mul::backend
