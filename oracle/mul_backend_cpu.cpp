// CPU product backend for the host library — TEST/BENCH INFRASTRUCTURE ONLY.
// Links the library's BigInt/BigFloat code to the threaded AVX2 restatement of the reference's
// NTT multiply (oracle/cpu_mt.cpp) so the π example can be run the reference's way on the host
// and its digits compared with the GPU run.  Never linked into bigint_amd/.
#include <algorithm>
#include <cstring>
#include <vector>

#include "../bigint_amd/host/mul_backend.h"

extern "C" int cpu_mul_digits(const uint32_t* a, size_t la, const uint32_t* b, size_t lb, int same,
                              uint32_t* out, size_t* nout);
extern "C" void cpu_init_thread_pool(uint32_t n);

namespace bigint::mul::backend {

void multiply(std::vector<Job>& jobs, int) {
    for (Job& j : jobs) {
        std::vector<uint32_t> tmp(j.a.size() + j.b.size() + 4);
        size_t n = 0;
        cpu_mul_digits(j.a.data(), j.a.size(), j.b.data(), j.b.size(), j.square ? 1 : 0, tmp.data(), &n);
        std::memcpy(j.out, tmp.data(), n * sizeof(uint32_t));
        j.out_len = n;
    }
}

auto multiply_top(std::span<const uint32_t> a, std::span<const uint32_t> b, bool square, size_t keep,
                  uint32_t* out) -> size_t {
    const size_t len = a.size() + b.size(), k = std::min(keep, len);
    std::vector<uint32_t> tmp(len + 4, 0u);
    size_t n = 0;
    cpu_mul_digits(a.data(), a.size(), b.data(), b.size(), square ? 1 : 0, tmp.data(), &n);
    std::memcpy(out, tmp.data() + (len - k), k * sizeof(uint32_t));
    return n;
}

bool load_planner(const char*) { return false; }  // the CPU restatement has no plans

}  // namespace bigint::mul::backend
