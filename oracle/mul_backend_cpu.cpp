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

// "Backend memory" of the CPU restatement is host memory.
auto dev_alloc(size_t words) -> std::shared_ptr<void> {
    return std::shared_ptr<void>(new uint32_t[words ? words : 1], [](void* p) { delete[] (uint32_t*)p; });
}

void dev_upload(uint32_t* dst, const std::vector<std::span<const uint32_t>>& src) {
    for (const auto& s : src) {
        std::memcpy(dst, s.data(), s.size() * 4);
        dst += s.size();
    }
}

void dev_download(const std::vector<std::span<uint32_t>>& out, const uint32_t* src) {
    for (const auto& o : out) {
        std::memcpy(o.data(), src, o.size() * 4);
        src += o.size();
    }
}

void dev_mul_batch(const std::vector<DevMul>& jobs, int) {
    for (const DevMul& j : jobs) {
        size_t la = j.la, lb = j.square ? j.la : j.lb;
        const uint32_t* b = j.square ? j.a : j.b;
        std::fill(j.out, j.out + la + lb, 0u);
        while (la > 1 && j.a[la - 1] == 0) --la;
        while (lb > 1 && b[lb - 1] == 0) --lb;
        std::vector<uint32_t> tmp(la + lb + 4, 0u);
        size_t n = 0;
        cpu_mul_digits(j.a, la, b, lb, (j.square && la == lb) ? 1 : 0, tmp.data(), &n);
        std::memcpy(j.out, tmp.data(), n * 4);
    }
}

void dev_add_batch(std::vector<DevAdd>& jobs) {
    for (DevAdd& j : jobs) {
        const size_t w = std::max(j.lx, j.ly);
        auto at = [](const uint32_t* v, size_t n, size_t i) { return i < n ? v[i] : 0u; };
        bool swap = false;
        if (j.subtract) {  // compare magnitudes from the top
            size_t i = w;
            while (i > 0 && at(j.x, j.lx, i - 1) == at(j.y, j.ly, i - 1)) --i;
            swap = i > 0 && at(j.y, j.ly, i - 1) > at(j.x, j.lx, i - 1);
        }
        const uint32_t* a = swap ? j.y : j.x;
        const uint32_t* b = swap ? j.x : j.y;
        const size_t la = swap ? j.ly : j.lx, lb = swap ? j.lx : j.ly;
        uint32_t c = 0;
        for (size_t i = 0; i < w; ++i) {
            const uint32_t x = at(a, la, i), y = at(b, lb, i);
            if (!j.subtract) {
                const uint32_t s = x + y + c;
                j.out[i] = s & 0xFFFFFFFu;
                c = s >> 28;
            } else {
                const uint32_t t = y + c;
                j.out[i] = (x - t) & 0xFFFFFFFu;
                c = x < t;
            }
        }
        size_t n = w;
        if (!j.subtract) j.out[n++] = c;
        while (n > 0 && j.out[n - 1] == 0) --n;
        j.swapped = swap;
        j.out_len = n;
    }
}

}  // namespace bigint::mul::backend
