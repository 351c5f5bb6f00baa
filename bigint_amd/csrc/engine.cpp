// Host side of the NTT multiply engine: root tables, per-length plans, workspace and the C ABI
// declared in include/bigint_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <tuple>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/bigint_hip.h"
#include "add.h"
#include "ntt_plan.h"
#include "ntt_primes.h"
#include "planner.h"

// Scratch of bi_hip_mul_device_noalloc: residues of both operands (three planes of n words each)
// and the carry-scan status words of one product on a transform of up to 2^kNoAllocMaxLog
// points.  Each device holds its own copy (24 MiB) from the time the module is loaded until the
// process ends.  Host code uses the device address from hipGetSymbolAddress, never the name.
constexpr uint32_t kNoAllocMaxLog = 20;
constexpr size_t kNoAllocWords = ((size_t)6 << kNoAllocMaxLog) + 1024;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

namespace bi {
namespace {

constexpr uint32_t kPh[3] = {kP[0], kP[1], kP[2]};  // the device primes and generators
constexpr uint32_t kGh[3] = {kG[0], kG[1], kG[2]};
constexpr int kTileH = 4096;

// Output tiles of the Garner + carry kernel for a product of `limbs` limbs (one carry-scan status
// word each).
constexpr uint32_t carry_tiles(size_t limbs) {
    return (uint32_t)((limbs + kCarryTileLimbs - 1) / kCarryTileLimbs);
}
constexpr uint32_t kMaxLog = 26;

// Device allocations made while serving calls (workspaces, staging, descriptors, caller buffers).
// With BI_HIP_TEST_FAIL_ALLOC set in the environment when the engine initialises, every such
// allocation fails, so tests can force the out-of-memory paths (BI_HIP_ENOMEM and the host
// library's std::bad_alloc) from the host; set-up allocations (root tables, plans) are not
// affected.
bool g_fail_alloc = false;
hipError_t call_malloc(void** p, size_t bytes) {
    if (g_fail_alloc) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

uint32_t mulmod(uint64_t a, uint64_t b, uint32_t p) { return (uint32_t)((a * b) % p); }
uint32_t powmod(uint32_t b, uint64_t e, uint32_t p) {
    uint64_t r = 1, x = b % p;
    while (e) {
        if (e & 1) r = r * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (uint32_t)r;
}
uint32_t invmod(uint32_t a, uint32_t p) { return powmod(a, p - 2, p); }  // p prime
uint32_t to_mont(uint32_t x, uint32_t p) { return (uint32_t)(((uint64_t)x << 32) % p); }
uint32_t bitrev(uint32_t x, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}
// primitive 2^k-th root of unity mod p (inverse root if inv)
uint32_t root(int I, uint32_t k, bool inv) {
    const uint32_t p = kPh[I];
    uint32_t w = powmod(kGh[I], (uint64_t)(p - 1) >> k, p);
    return inv ? invmod(w, p) : w;
}

struct OuterPass {
    uint32_t log_r, log_c, log_s;
    uint32_t nblk;
    uint32_t* beta = nullptr;      // device [3][nblk] (null when nblk == 1), both directions
};

struct Plan {
    uint32_t log_n = 0;
    std::vector<OuterPass> outer;  // forward order
    uint32_t mid_log_r = 0, mid_nblk = 1;
    uint32_t* mid_beta = nullptr;
    uint32_t kscale[3] = {0, 0, 0};
};

// Per-block roots of a pass whose tiles have 2^log_r rows: for block b,
// beta_b = w_{n/S}^{bitrev_{log nblk}(b)}, stored as beta_b^(2^e) for e < log_r (Montgomery form,
// kBetaPowStride words per block), so a tile's twiddles tw[k] = W[k] * beta^tr need one product
// each and no squaring chain on the device.  The inverse stages use the same forward roots.
hipError_t upload_beta(uint32_t log_n, uint32_t log_r, uint32_t log_s, uint32_t nblk, uint32_t** fwd) {
    *fwd = nullptr;
    if (nblk <= 1) return hipSuccess;
    if (log_r > kBetaPowStride) return hipErrorInvalidValue;
    const size_t stride = kBetaPowStride;
    std::vector<uint32_t> hf(3 * (size_t)nblk * stride, 0u);
    const uint32_t lb = 31u - (uint32_t)__builtin_clz(nblk);
    for (int I = 0; I < 3; ++I) {
        const uint32_t p = kPh[I];
        const uint32_t wf = root(I, log_n - log_s, false);
        for (uint32_t b = 0; b < nblk; ++b) {
            uint32_t bf = powmod(wf, bitrev(b, lb), p);
            uint32_t* of = &hf[((size_t)I * nblk + b) * stride];
            for (uint32_t k = 0; k < log_r; ++k) {
                of[k] = to_mont(bf, p);
                bf = mulmod(bf, bf, p);
            }
        }
    }
    hipError_t e = hipMalloc((void**)fwd, hf.size() * 4);
    if (e != hipSuccess) return e;
    return hipMemcpy(*fwd, hf.data(), hf.size() * 4, hipMemcpyHostToDevice);
}

struct ProfRec {
    int kind;
    hipEvent_t e0, e1;
    double bytes;
};

// Host <-> device copies of caller (pageable) memory through two pinned chunks: the host copy of
// one chunk overlaps the DMA of the other, and large host copies are split over a few threads.
// Host side of a transfer: byte ranges in host memory, concatenated in order.
struct HostSegs {
    std::vector<char*> ptr;
    std::vector<size_t> start;  // offset of each range in the concatenation, plus the total
    HostSegs() : start{0} {}
    void add(const void* p, size_t bytes) {
        if (!bytes) return;
        ptr.push_back((char*)p);
        start.push_back(start.back() + bytes);
    }
    size_t total() const { return start.back(); }
    // bytes [off, off + n) of the concatenation to (to_buf) or from buf
    void copy(size_t off, size_t n, char* buf, bool to_buf) const {
        size_t i = (size_t)(std::upper_bound(start.begin(), start.end(), off) - start.begin()) - 1;
        while (n) {
            const size_t in = off - start[i], len = std::min(n, start[i + 1] - off);
            if (to_buf) std::memcpy(buf, ptr[i] + in, len);
            else std::memcpy(ptr[i] + in, buf, len);
            buf += len;
            off += len;
            n -= len;
            ++i;
        }
    }
};

// A few persistent threads for the host side of staged copies: f(i), i < n, runs on them and on
// the calling thread.  (Starting threads per chunk cost about 10 us each and limited the split to
// four; one copy thread moves about 31 GB/s on the GPU box's host, eight about 75 GB/s.)
class CopyWorkers {
  public:
    void run(unsigned n, const std::function<void(unsigned)>& f) {
        if (n <= 1) {
            if (n) f(0);
            return;
        }
        std::unique_lock<std::mutex> lk(m_);
        start();
        f_ = &f;
        n_ = n;
        next_ = 0;
        left_ = n;
        ++gen_;
        cv_.notify_all();
        work(lk);
        done_.wait(lk, [&] { return left_ == 0; });
        f_ = nullptr;
    }
    static unsigned threads() {
        const unsigned hc = std::thread::hardware_concurrency();
        return std::max(1u, std::min(8u, hc / 2));
    }

  private:
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(unsigned)>* f_ = nullptr;
    unsigned n_ = 0, next_ = 0, left_ = 0;
    uint64_t gen_ = 0;
    bool started_ = false;

    void start() {  // under m_
        if (started_) return;
        started_ = true;
        for (unsigned t = 1; t < threads(); ++t) std::thread([this] { loop(); }).detach();
    }
    // take tasks of the current generation until none is left (called with m_ held)
    void work(std::unique_lock<std::mutex>& lk) {
        while (f_ && next_ < n_) {
            const unsigned i = next_++;
            const auto* f = f_;
            lk.unlock();
            (*f)(i);
            lk.lock();
            if (--left_ == 0) done_.notify_all();
        }
    }
    void loop() {
        std::unique_lock<std::mutex> lk(m_);
        uint64_t seen = 0;
        for (;;) {
            cv_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            work(lk);
        }
    }
};

// The copy workers live until the process ends (never destroyed: detached threads wait on them).
CopyWorkers& copy_workers() {
    static CopyWorkers* w = new CopyWorkers();
    return *w;
}

// Host <-> device copies through two pinned buffers: the host-side copy of one chunk (split over
// the copy workers) overlaps the DMA of the other.  The host side may be a list of ranges, so
// batches gather operands and scatter products without a staging copy.
class PinnedPipe {
  public:
    static constexpr size_t kChunk = size_t(16) << 20;  // bytes per pinned buffer

    ~PinnedPipe() = default;  // pinned buffers live until the process ends

    hipError_t h2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
        HostSegs h;
        h.add(src, bytes);
        return h2d(dst, h, st);
    }
    hipError_t d2h(void* dst, const void* src, size_t bytes, hipStream_t st) {
        HostSegs h;
        h.add(dst, bytes);
        return d2h(h, src, st);
    }

    // Device bytes [0, src.total()) at dst receive the concatenation of src.
    hipError_t h2d(void* dst, const HostSegs& src, hipStream_t st) {
        hipError_t e = ensure();
        const size_t bytes = src.total();
        for (size_t off = 0; e == hipSuccess && off < bytes; off += kChunk) {
            const size_t n = std::min(kChunk, bytes - off), b = next_;
            next_ ^= 1;
            // the buffer may still feed a DMA of this call or of an earlier one
            if (pending_[b]) e = hipEventSynchronize(ev_[b]);
            if (e != hipSuccess) break;
            pending_[b] = false;
            copy(src, off, n, (char*)buf_[b], true);
            e = hipMemcpyAsync((char*)dst + off, buf_[b], n, hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = hipEventRecord(ev_[b], st);
            pending_[b] = e == hipSuccess;
        }
        return e;
    }

    // Device bytes [0, dst.total()) at src go to the ranges of dst.  Waits for earlier work on
    // st; returns when dst holds the data.
    hipError_t d2h(const HostSegs& dst, const void* src, hipStream_t st) {
        hipError_t e = ensure();
        const size_t bytes = dst.total();
        const size_t nchunks = (bytes + kChunk - 1) / kChunk;
        auto issue = [&](size_t k) {
            const size_t off = k * kChunk, n = std::min(kChunk, bytes - off), b = k & 1;
            hipError_t r = hipMemcpyAsync(buf_[b], (const char*)src + off, n, hipMemcpyDeviceToHost, st);
            return r == hipSuccess ? hipEventRecord(ev_[b], st) : r;
        };
        // DMAs into a buffer are ordered after any pending DMA out of it on the same stream
        if (e == hipSuccess && nchunks) e = issue(0);
        for (size_t k = 0; e == hipSuccess && k < nchunks; ++k) {
            if (k + 1 < nchunks) e = issue(k + 1);  // the other buffer was emptied in step k-1
            if (e == hipSuccess) e = hipEventSynchronize(ev_[k & 1]);
            if (e != hipSuccess) break;
            const size_t off = k * kChunk;
            copy(dst, off, std::min(kChunk, bytes - off), (char*)buf_[k & 1], false);
        }
        pending_[0] = pending_[1] = false;  // every event recorded above has completed
        return e;
    }

  private:
    void* buf_[2] = {nullptr, nullptr};
    hipEvent_t ev_[2] = {nullptr, nullptr};
    bool pending_[2] = {false, false};  // ev_[i] marks a DMA out of buf_[i] not yet waited for
    size_t next_ = 0;                   // buffer for the next host-to-device chunk

    hipError_t ensure() {
        for (int i = 0; i < 2; ++i) {
            if (!buf_[i]) {
                const hipError_t e = hipHostMalloc(&buf_[i], kChunk, hipHostMallocDefault);
                if (e != hipSuccess) return e;
            }
            if (!ev_[i]) {
                const hipError_t e = hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming);
                if (e != hipSuccess) return e;
            }
        }
        return hipSuccess;
    }

    // bytes [off, off + n) of h to or from buf, in pieces of at least 1 MiB over the copy workers
    static void copy(const HostSegs& h, size_t off, size_t n, char* buf, bool to_buf) {
        constexpr size_t kPiece = size_t(1) << 20;
        const unsigned np = (unsigned)std::min<size_t>(CopyWorkers::threads(), std::max<size_t>(1, n / kPiece));
        const size_t per = (n / np + 63) & ~size_t(63);
        copy_workers().run(np, [&](unsigned t) {
            const size_t lo = std::min(n, t * per), hi = t + 1 == np ? n : std::min(n, (t + 1) * per);
            if (hi > lo) h.copy(off + lo, hi - lo, buf + lo, to_buf);
        });
    }
};

class Engine {
public:
    std::mutex mtx;
    uint32_t prof_mask = 0;  // bit k: bracket launches of kernel class k with events
    std::vector<ProfRec> prof;
    std::vector<hipEvent_t> event_pool;

    hipEvent_t get_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    // Bracket one launch with events when its kernel class is being profiled.
    bool suppress_prof = false;  // set while the allocation-free entry point launches
    template <class F>
    hipError_t timed(int kind, double bytes, hipStream_t st, F&& launch) {
        if (suppress_prof || !((prof_mask >> kind) & 1u)) return launch();
        ProfRec r{kind, get_event(), get_event(), bytes};
        if (!r.e0 || !r.e1) return hipErrorOutOfMemory;
        hipError_t e = hipEventRecord(r.e0, st);
        if (e == hipSuccess) e = launch();
        if (e == hipSuccess) e = hipEventRecord(r.e1, st);
        if (e != hipSuccess) {
            event_pool.push_back(r.e0);
            event_pool.push_back(r.e1);
            return e;
        }
        prof.push_back(r);
        return e;
    }
    void prof_clear() {
        for (auto& r : prof) {
            event_pool.push_back(r.e0);
            event_pool.push_back(r.e1);
        }
        prof.clear();
    }
    bool ready = false;
    int device = -1;
    hipStream_t stream = nullptr;
    // Plan inference (the MLP kernels) runs on a stream of its own: the caller waits for the
    // choices, and on the caller's stream that wait would also cover every product queued there.
    hipStream_t plan_stream = nullptr;
    uint32_t* wtab = nullptr;  // [3][4096]
    GarnerConsts gk{};
    std::map<uint32_t, std::unique_ptr<Plan>> plans;
    // workspace (grow-only)
    uint32_t* ws = nullptr;
    size_t ws_words = 0;
    // staging for host-pointer calls
    uint32_t* io = nullptr;
    size_t io_words = 0;

    int init(int dev) {
        if (ready) return BI_HIP_OK;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return BI_HIP_ENODEV;
        if (dev < 0 || dev >= count) return BI_HIP_EINVAL;
        if (hipSetDevice(dev) != hipSuccess) return BI_HIP_EDEVICE;
        device = dev;
        g_fail_alloc = std::getenv("BI_HIP_TEST_FAIL_ALLOC") != nullptr;
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&plan_stream, hipStreamNonBlocking) != hipSuccess)
            return BI_HIP_EDEVICE;
        // root table W[m + i] = w_{2m}^{bitrev_{log m}(i)}, Montgomery form, one per prime; the
        // inverse stages use it too (their output order is mirrored, which Garner undoes)
        std::vector<uint32_t> h(3 * (size_t)kTileH);
        for (int I = 0; I < 3; ++I) {
            uint32_t* t = h.data() + (size_t)I * kTileH;
            t[0] = to_mont(1, kPh[I]);
            for (uint32_t lm = 0; (1u << lm) < (uint32_t)kTileH; ++lm) {
                const uint32_t m = 1u << lm;
                const uint32_t w = root(I, lm + 1, false);
                for (uint32_t i = 0; i < m; ++i)
                    t[m + i] = to_mont(powmod(w, bitrev(i, lm), kPh[I]), kPh[I]);
            }
        }
        if (hipMalloc((void**)&wtab, h.size() * 4) != hipSuccess) return BI_HIP_ENOMEM;
        if (hipMemcpy(wtab, h.data(), h.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            return BI_HIP_EDEVICE;
        const uint32_t p0 = kPh[0], p1 = kPh[1], p2 = kPh[2];
        gk.inv01 = to_mont(invmod(p0 % p1, p1), p1);
        gk.inv02 = to_mont(invmod(p0 % p2, p2), p2);
        gk.inv12 = to_mont(invmod(p1 % p2, p2), p2);
        // set-up of the allocation-free entry point on this device: the scratch array's device
        // address, and every plan it may use built now (plan() allocates its root tables)
        if (hipGetSymbolAddress((void**)&noalloc_scratch, HIP_SYMBOL(::scratch)) != hipSuccess || !noalloc_scratch)
            return BI_HIP_EDEVICE;
        for (uint32_t lg = 4; lg <= kNoAllocMaxLog; ++lg)
            if (!plan(lg)) return BI_HIP_ENOMEM;
        ready = true;
        return BI_HIP_OK;
    }

    uint32_t* noalloc_scratch = nullptr;  // device address of `scratch` on this engine's device

    // An existing plan, never built here (so nothing is allocated).
    const Plan* find_plan(uint32_t log_n, uint32_t mid_lr = 12) const {
        if (log_n <= mid_lr) mid_lr = log_n;
        const auto it = plans.find(log_n * 32u + mid_lr);
        return it == plans.end() ? nullptr : it->second.get();
    }

    // Stage split for a transform of 2^log_n points whose contiguous block pass has 2^mid_lr
    // points: the remaining outer stages go in one pass (<= 8 stages) or two.  Returns false when
    // a pass would have more tile columns than its row stride allows.
    static bool split_outer(uint32_t log_n, uint32_t mid_lr, std::vector<uint32_t>& rs) {
        rs.clear();
        if (log_n <= mid_lr) return true;
        const uint32_t q = log_n - mid_lr;
        if (q <= 8) rs = {q};
        else if (q <= 16) rs = {(q + 1) / 2, q / 2};
        else return false;
        uint32_t done = 0;
        for (uint32_t r : rs) {
            const uint32_t log_s = log_n - done - r;
            if (12 - r > log_s) return false;
            done += r;
        }
        return true;
    }

    Plan* plan(uint32_t log_n, uint32_t mid_lr = 12) {
        if (log_n <= mid_lr) mid_lr = log_n;
        const uint32_t key = log_n * 32u + mid_lr;
        auto it = plans.find(key);
        if (it != plans.end()) return it->second.get();
        std::vector<uint32_t> rs;
        if (mid_lr < 4 || mid_lr > 12 || !split_outer(log_n, mid_lr, rs)) return nullptr;
        auto pl = std::make_unique<Plan>();
        pl->log_n = log_n;
        const uint32_t n = 1u << log_n;
        uint32_t done = 0;
        for (uint32_t r : rs) {
            OuterPass op;
            op.log_r = r;
            op.log_c = 12 - r;
            op.log_s = log_n - done - r;
            op.nblk = 1u << done;
            if (upload_beta(log_n, r, op.log_s, op.nblk, &op.beta) != hipSuccess)
                return nullptr;
            pl->outer.push_back(op);
            done += r;
        }
        pl->mid_log_r = mid_lr;
        pl->mid_nblk = n >> mid_lr;
        if (upload_beta(log_n, mid_lr, 0, pl->mid_nblk, &pl->mid_beta) != hipSuccess)
            return nullptr;
        for (int I = 0; I < 3; ++I) {
            const uint32_t p = kPh[I];
            const uint32_t r2 = mulmod(to_mont(1, p), to_mont(1, p), p);  // R^2 mod p
            pl->kscale[I] = mulmod(invmod(n % p, p), r2, p);
        }
        Plan* raw = pl.get();
        plans[key] = std::move(pl);
        return raw;
    }

    // The workspace, its status words and the descriptor buffer are shared by every call, whatever
    // stream it runs on.  shared_done is recorded after the last queued use (on shared_stream), so
    // a call on another stream first waits for it: work queued on two streams without a host
    // synchronisation between them still uses the buffers one call after the other.
    hipEvent_t shared_done = nullptr;
    hipStream_t shared_stream = nullptr;

    // Queue, on st, a wait for the last use of the shared buffers on another stream.
    int order_shared(hipStream_t st) {
        if (!shared_stream || shared_stream == st) return BI_HIP_OK;
        return hipStreamWaitEvent(st, shared_done, 0) == hipSuccess ? BI_HIP_OK : BI_HIP_EDEVICE;
    }
    // Block the host until every queued use of the shared buffers is done (before it rewrites the
    // descriptor buffer with a synchronous copy).
    int wait_shared() {
        if (!shared_stream) return BI_HIP_OK;
        return hipEventSynchronize(shared_done) == hipSuccess ? BI_HIP_OK : BI_HIP_EDEVICE;
    }
    // Record that work using the shared buffers has been queued on st.
    int used_shared(hipStream_t st) {
        if (!shared_done && hipEventCreateWithFlags(&shared_done, hipEventDisableTiming) != hipSuccess) {
            shared_done = nullptr;
            return BI_HIP_EDEVICE;
        }
        if (hipEventRecord(shared_done, st) != hipSuccess) return BI_HIP_EDEVICE;
        shared_stream = st;
        return BI_HIP_OK;
    }

    // device workspace of at least `words` u32
    uint32_t* workspace(size_t words) {
        if (words > ws_words) {
            // queued kernels may still use the old buffer: let them finish before it goes
            if (ws && hipDeviceSynchronize() != hipSuccess) return nullptr;
            graph_reset();
            if (ws) (void)hipFree(ws);
            ws = nullptr;
            ws_words = 0;
            if (call_malloc((void**)&ws, words * 4) != hipSuccess) return nullptr;
            ws_words = words;
        }
        return ws;
    }

    enum Kind { kPlain = 0, kSquareKind = 1, kBForward = 2, kChunk = 3, kInverse = 4 };

    // Launch the pass sequence of one plan for njobs jobs of the same transform length.
    //   kPlain / kSquareKind: product of A and B (or A^2), Garner+carry into the output;
    //   kBForward: forward transform of operand A only, left in res_a (no Garner);
    //   kChunk: A-chunk transform, pointwise with the shared transformed B in res_b, inverse,
    //           Garner summing the chunk products.
    // single/gsingle are used when njobs == 1 and no descriptor arrays are given.
    int launch_products(const Plan* pl, Kind kind, uint32_t njobs, const PassJob& single,
                        const PassJob* d_jobs, uint32_t ngjobs, const GarnerJob& gsingle,
                        const GarnerJob* d_gjobs, uint32_t max_out_len, double in_limbs,
                        double out_limbs, uint32_t* status, uint32_t status_words, hipStream_t st) {
        const uint32_t n = 1u << pl->log_n;
        const double nj = (double)njobs;
        // the first kernel of the sequence clears the Garner launch's status words
        bool clear_pending = kind != kBForward && status && status_words;
        auto take_clear = [&](PassArgs& x) {
            x.clear = clear_pending ? status : nullptr;
            x.clear_words = clear_pending ? status_words : 0u;
            clear_pending = false;
        };
        const bool one_op = kind != kPlain;
        const double nops_d = one_op ? 1.0 : 2.0;
        const bool fwd_passes = kind != kInverse;  // kInverse: res_a already holds a spectrum
        PassArgs a{};
        a.job = single;
        a.jobs = d_jobs;
        a.log_n = pl->log_n;
        a.wtab = wtab;
        const uint32_t base_flags = kind == kSquareKind ? (uint32_t)kSquare
                                  : (kind == kPlain ? 0u : (uint32_t)kSingle);
        const uint32_t nops = one_op ? 1u : 2u;
        for (size_t k = 0; fwd_passes && k < pl->outer.size(); ++k) {
            const OuterPass& op = pl->outer[k];
            a.log_r = op.log_r;
            a.log_c = op.log_c;
            a.log_s = op.log_s;
            a.flags = base_flags | (k == 0 ? (uint32_t)kFromLimbs : 0u);
            a.beta = op.beta;
            a.beta_stride = op.nblk;
            take_clear(a);
            // the first pass reads the limbs once; every pass writes 3 planes per operand
            const double bytes = k == 0 ? 4.0 * in_limbs + 12.0 * n * nops_d * nj : 24.0 * n * nops_d * nj;
            if (timed(k == 0 ? BI_HIP_K_OUTER_FIRST : BI_HIP_K_OUTER_FWD, bytes, st,
                      [&] { return launch_outer_pass(a, false, njobs, nops, st); }) != hipSuccess)
                return BI_HIP_EDEVICE;
        }
        a.log_r = pl->mid_log_r;
        a.log_c = 0;
        a.log_s = 0;
        a.flags = (kind == kSquareKind ? (uint32_t)kSquare : 0u) |
                  (kind == kBForward ? (uint32_t)kFwdOnly : 0u) |
                  (kind == kChunk ? (uint32_t)kBPre : 0u) |
                  (kind == kInverse ? (uint32_t)kInvOnly : 0u) |
                  (pl->outer.empty() && fwd_passes ? (uint32_t)kFromLimbs : 0u);
        a.beta = pl->mid_beta;
        a.beta_stride = pl->mid_nblk;
        take_clear(a);
        {
            double in_bytes = pl->outer.empty() && fwd_passes ? 4.0 * in_limbs : 12.0 * n * nops_d * nj;
            if (kind == kChunk) in_bytes += 12.0 * n * nj;  // the shared transformed B
            if (timed(BI_HIP_K_MID, in_bytes + 12.0 * n * nj, st,
                      [&] { return launch_mid_pass(a, njobs, st); }) != hipSuccess)
                return BI_HIP_EDEVICE;
        }
        if (kind == kBForward) return BI_HIP_OK;
        for (size_t k = pl->outer.size(); k-- > 0;) {
            const OuterPass& op = pl->outer[k];
            a.log_r = op.log_r;
            a.log_c = op.log_c;
            a.log_s = op.log_s;
            a.flags = 0;
            a.beta = op.beta;  // forward roots (see coeff_pos in the kernels)
            a.beta_stride = op.nblk;
            take_clear(a);
            if (timed(BI_HIP_K_OUTER_INV, 24.0 * n * nj, st,
                      [&] { return launch_outer_pass(a, true, njobs, 1, st); }) != hipSuccess)
                return BI_HIP_EDEVICE;
        }
        GarnerArgs g{};
        g.job = gsingle;
        g.jobs = d_gjobs;
        g.k = gk;
        if (timed(BI_HIP_K_GARNER, 12.0 * n * nj + 4.0 * out_limbs, st,
                  [&] {
                      return launch_garner_carry(g, max_out_len, ngjobs, kind == kChunk || kind == kInverse, st);
                  }) != hipSuccess)
            return BI_HIP_EDEVICE;
        return BI_HIP_OK;
    }

    // One product.  planned: the engine's plan choice (table, else heuristic), which may chunk or
    // re-split; otherwise plan 0 (one unchunked transform with a 4096-point block pass).
    int run(const uint32_t* d_a, uint32_t la, const uint32_t* d_b, uint32_t lb, bool square,
            uint32_t* d_out, hipStream_t st, bool planned = true) {
        const int lg0 = bi_hip_transform_log2(la, lb);
        if (!product_supported(la, lb)) return BI_HIP_ETOOBIG;
        if (needs_pieces(la, lb)) return run_pieces(d_a, la, d_b, lb, d_out, st);
        // the table (or the heuristic) may pick a chunked or re-split plan, and products beyond
        // one transform need one: those run through the batch path
        const bi_hip_dpair p{d_a, la, square ? nullptr : d_b, lb, d_out, (size_t)la + lb};
        if (lg0 < 0) return run_batch(&p, 1, BI_HIP_PLAN_TABLE, nullptr, st);
        std::vector<PairPlan> chosen;
        if (planned) {
            const int rc = choose_plans(&p, 1, BI_HIP_PLAN_TABLE, nullptr, chosen, st);
            if (rc) return rc;
        }
        // chunked and re-split plans need the batch path's shared buffers; an unchunked plan with
        // any block-pass size launches directly (kernel arguments by value: no descriptor upload,
        // no wait for earlier work on the stream)
        if (planned && (chosen[0].chunk_len || chosen[0].split_both || chosen[0].lg != (uint32_t)lg0))
            return run_batch(&p, 1, BI_HIP_PLAN_TABLE, nullptr, st);
        const uint32_t lg = (uint32_t)lg0;
        Plan* pl = planned ? plan(lg, chosen[0].mid_lr) : plan(lg);
        if (!pl) return BI_HIP_ENOMEM;
        uint32_t* w = workspace(single_words(lg, la, lb, square));
        if (!w) return BI_HIP_ENOMEM;
        if (int rc = order_shared(st)) return rc;
        if (int rc = launch_single(pl, d_a, la, d_b, lb, square, d_out, w, st)) return rc;
        return used_shared(st);
    }

    // Device words one unchunked product needs: the residues of both operands (one for a square)
    // and the carry-scan status words.
    static size_t single_words(uint32_t lg, uint32_t la, uint32_t lb, bool square) {
        return ((size_t)3 << lg) * (square ? 1 : 2) + carry_tiles((size_t)la + lb) + 64;
    }

    // Queue one unchunked product on plan pl with its residues and status words in w (at least
    // single_words() words): no allocation and no host synchronisation.
    int launch_single(const Plan* pl, const uint32_t* d_a, uint32_t la, const uint32_t* d_b, uint32_t lb,
                      bool square, uint32_t* d_out, uint32_t* w, hipStream_t st) {
        const uint32_t n = 1u << pl->log_n;
        const uint32_t out_len = la + lb;
        const uint32_t ntiles = carry_tiles(out_len);
        uint32_t* res_a = w;
        uint32_t* res_b = square ? nullptr : w + (size_t)3 * n;
        uint32_t* status = w + (size_t)3 * n * (square ? 1 : 2);
        const PassJob job{d_a, square ? d_a : d_b, la, square ? la : lb, res_a, res_b, 0u, ~0u};
        GarnerJob gj{};
        gj.y = res_a;
        gj.out = d_out;
        gj.n = n;
        gj.out_len = out_len;
        for (int I = 0; I < 3; ++I) gj.kscale[I] = pl->kscale[I];
        gj.status = status;
        gj.fix = status + ntiles;
        if ((uint64_t)words_of(la) + words_of(lb) - 1 > n) {
            gj.wrap_a = d_a;
            gj.wrap_b = square ? d_a : d_b;
            gj.wrap_la = la;
            gj.wrap_lb = lb;
        }
        return launch_products(pl, square ? kSquareKind : kPlain, 1, job, nullptr, 1, gj, nullptr, out_len,
                               (double)la + (square ? 0.0 : (double)lb), (double)out_len, status,
                               ntiles + 1, st);
    }

    // Batches as captured graphs: on by default (bi_hip_set_graphs).  Two executable graphs are
    // kept; each is updated in place when a later batch has its shape.  A batch call does not wait
    // for the previous one: it takes the slot launched last if that launch is done (a caller that
    // synchronises between batches keeps updating one graph), else the other slot, whose launch
    // it waits for (its `done` event) before the update; batches queued back to back use the two
    // in turn.
    struct GraphSlot {
        hipGraphExec_t exec = nullptr;
        hipEvent_t done = nullptr;  // recorded after the slot's last launch
        bool launched = false;
    };
    bool graphs = true;
    GraphSlot gslots[2];
    int gslot_last = 0;

    // Drop the executable graphs: called before any buffer they may reference is freed.
    void graph_reset() {
        for (GraphSlot& g : gslots) {
            if (g.launched) (void)hipEventSynchronize(g.done);
            g.launched = false;
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
        }
    }

    template <class F>
    int launch_as_graph(F&& launch_all, hipStream_t st) {
        int si = gslot_last;
        if (gslots[si].launched) {
            const hipError_t q = hipEventQuery(gslots[si].done);
            if (q == hipErrorNotReady) {
                (void)hipGetLastError();  // a query of a pending event is not an error
                si ^= 1;
            } else if (q != hipSuccess) {
                return BI_HIP_EDEVICE;
            }
        }
        GraphSlot& slot = gslots[si];
        gslot_last = si;
        if (!slot.done && hipEventCreateWithFlags(&slot.done, hipEventDisableTiming) != hipSuccess) {
            slot.done = nullptr;
            return BI_HIP_EDEVICE;
        }
        // the slot's previous launch must be done before its executable graph changes
        if (slot.launched && hipEventSynchronize(slot.done) != hipSuccess) return BI_HIP_EDEVICE;
        slot.launched = false;
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return BI_HIP_EDEVICE;
        const int rc = launch_all();
        hipGraph_t g = nullptr;
        const hipError_t end = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (end != hipSuccess || !g) return BI_HIP_EDEVICE;
        bool ok = false;
        if (slot.exec) {
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult res{};
            ok = hipGraphExecUpdate(slot.exec, g, &err_node, &res) == hipSuccess &&
                 res == hipGraphExecUpdateSuccess;
            if (!ok) {  // another shape: a new executable graph (and no stale error left behind)
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(slot.exec);
                slot.exec = nullptr;
            }
        }
        if (!ok && hipGraphInstantiate(&slot.exec, g, nullptr, nullptr, 0) != hipSuccess) {
            slot.exec = nullptr;
            (void)hipGraphDestroy(g);
            return BI_HIP_EDEVICE;
        }
        (void)hipGraphDestroy(g);
        if (hipGraphLaunch(slot.exec, st) != hipSuccess || hipEventRecord(slot.done, st) != hipSuccess)
            return BI_HIP_EDEVICE;
        slot.launched = true;
        return BI_HIP_OK;
    }

    // Pinned host staging of the batch descriptors, two buffers used in turn: the descriptors go
    // to the device with a copy ordered on the caller's stream after the earlier work that reads
    // the device buffer, so the host does not wait for that work; a staging buffer is rewritten
    // only once its copy of two calls earlier is done.
    struct DescStage {
        uint8_t* host = nullptr;
        size_t bytes = 0;
        hipEvent_t copied = nullptr;
        bool pending = false;
    };
    DescStage dstage[2];
    int dstage_next = 0;

    // Queue the copy of `bytes` descriptor bytes from src into the device buffer on st (ordered
    // after every earlier use of the shared buffers).
    int upload_desc(const uint8_t* src, size_t bytes, hipStream_t st) {
        if (bytes == 0) return order_shared(st);  // no grouped launches (products in pieces only)
        DescStage& d = dstage[dstage_next];
        dstage_next ^= 1;
        if (!d.copied && hipEventCreateWithFlags(&d.copied, hipEventDisableTiming) != hipSuccess) {
            d.copied = nullptr;
            return BI_HIP_EDEVICE;
        }
        if (d.pending && hipEventSynchronize(d.copied) != hipSuccess) return BI_HIP_EDEVICE;
        d.pending = false;
        if (bytes > d.bytes) {
            if (d.host) (void)hipHostFree(d.host);
            d.host = nullptr;
            d.bytes = 0;
            if (hipHostMalloc((void**)&d.host, bytes, hipHostMallocDefault) != hipSuccess) {
                d.host = nullptr;
                return BI_HIP_ENOMEM;
            }
            d.bytes = bytes;
        }
        std::memcpy(d.host, src, bytes);
        if (int rc = order_shared(st)) return rc;
        if (hipMemcpyAsync(desc, d.host, bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipEventRecord(d.copied, st) != hipSuccess)
            return BI_HIP_EDEVICE;
        d.pending = true;
        return BI_HIP_OK;
    }

    // descriptor buffer for batched launches (device), grow-only
    uint8_t* desc = nullptr;
    size_t desc_bytes = 0;
    PinnedPipe pipe;  // host <-> device copies of the host-pointer entry points
    std::mutex pipe_mtx;  // taken after mtx when both are held
    // device staging of bi_hip_mul_batch (host operands in, products out), grow-only
    std::mutex stage_mtx;
    uint32_t* stage = nullptr;
    size_t stage_words = 0;
    // scratch of bi_hip_add_batch_device (descriptors, tile maps, carries, flags), grow-only
    std::mutex add_mtx;
    uint8_t* add_scratch = nullptr;
    size_t add_bytes = 0;

    // Plan of one product: transform length, optional chunking of the longer operand, and the
    // size of the contiguous block pass (which fixes the radix split of the outer stages).
    // A chunked product transforms the shorter operand once and reuses it for every chunk.
    struct PairPlan {
        uint32_t lg;         // log2 of the transform length
        uint32_t chunk_len;  // 0: ordinary product; else words per chunk of the longer operand
        uint32_t mid_lr;     // log2 of the block-pass size
        bool split_both = false;  // both operands cut into kSplitChunk-word chunks (see below)
    };

    // Products beyond one 2^26-point transform whose shorter operand is long too: both operands
    // are cut into 2^25-word chunks, every chunk is transformed once, the pointwise products of
    // chunk pairs are summed per diagonal (i + k = s) in the transform domain, and each diagonal
    // is inverse-transformed once (the reference switches to its SSA tier there, mul.h:199-205).
    // Coefficients stay below min(wa, wb) 2^64 <= 2^90 < P0 P1 P2 while min(wa, wb) <= 2^26.
    static constexpr uint32_t kSplitLog = 25;
    static constexpr uint32_t kSplitChunk = 1u << kSplitLog;
    // The product's limb count must stay below 2^32 - 2^13, so tile counts such as
    // carry_tiles(la + lb) and word offsets stay exact in the kernels' 32-bit arithmetic.
    static constexpr uint64_t kMaxProductLimbs = 0xFFFFFFFFull - 8192u;
    static bool product_supported(size_t la, size_t lb) {
        if (la == 0 || lb == 0 || la > kMaxProductLimbs || lb > kMaxProductLimbs) return false;
        return la + lb <= kMaxProductLimbs;
    }
    // Products whose shorter operand has more than 2^26 words are beyond one set of diagonal sums
    // (their coefficients would pass P0 P1 P2): the shorter operand is cut into limb-aligned
    // pieces of kPieceLimbs limbs (2^26 - 4 words), each piece's product with the longer operand
    // runs as a product of its own, and the partial products are added into the output at their
    // limb offsets (run_pieces; the reference's SSA tier takes these sizes, mul.h:199-205).
    static constexpr uint32_t kPieceLimbs = 76695840u;  // 8 * 9586980: words_of() = 67108860
    static bool needs_pieces(size_t la, size_t lb) {
        return std::min(words_of((uint32_t)la), words_of((uint32_t)lb)) > (1u << 26);
    }
    static constexpr int kNumPlanIds = 8;     // the plan space of the table and the MLP
    static constexpr int kSplitPlanId = 8;    // both operands chunked (explicit, or beyond the rest)

    static uint32_t ceil_log2(uint64_t x) {
        uint32_t c = 0;
        while (((uint64_t)1 << c) < x) ++c;
        return c;
    }

    // Plan ids: 0/1/2 = unchunked with 4096/2048/1024-point block pass; 3/4/5 = chunks of
    // 2^c, 2^(c+1), 2^(c+2) limbs, 2^c the smallest power of two >= the shorter operand;
    // 6/7 = the longer operand split into 2 / 3 chunks (multiples of 4096 limbs).  Chunk plans
    // are allowed only where their transform is shorter than the unchunked one.
    static bool plan_for_id(uint32_t la, uint32_t lb, bool sq, int id, PairPlan& out) {
        const int lg0 = bi_hip_transform_log2(la, lb);  // -1: beyond one transform
        if (id == kSplitPlanId && product_supported(la, lb) && !needs_pieces(la, lb)) {
            out = PairPlan{26, kSplitChunk, 12, true};
            return true;
        }
        if (id < 0 || id >= kNumPlanIds || !product_supported(la, lb)) return false;
        std::vector<uint32_t> rs;
        if (id <= 2) {
            if (lg0 < 0) return false;
            const uint32_t mid = 12u - (uint32_t)id;
            if (id > 0 && (uint32_t)lg0 <= mid) return false;  // same as plan 0
            if (!split_outer((uint32_t)lg0, mid, rs)) return false;
            out = PairPlan{(uint32_t)lg0, 0, mid};
            return true;
        }
        if (sq) return false;
        // chunking works on the operands' 32-bit transform words
        const uint32_t hi = words_of(std::max(la, lb)), lo = words_of(std::min(la, lb));
        uint32_t L;
        if (id <= 5) {
            uint32_t c = std::max<uint32_t>(10, ceil_log2(lo)) + (uint32_t)(id - 3);
            if (c >= 31) return false;
            L = 1u << c;
        } else {
            const uint32_t parts = (uint32_t)id - 4;  // 2 or 3
            L = (hi + parts - 1) / parts;
            L = (L + kCarryTileWords - 1) / kCarryTileWords * kCarryTileWords;
        }
        if (L >= hi) return false;
        const uint32_t lg = std::max<uint32_t>(4, ceil_log2((uint64_t)L + lo - 1));
        if ((lg0 >= 0 && lg >= (uint32_t)lg0) || lg > 26) return false;
        const uint32_t mid = std::min<uint32_t>(12, lg);
        if (!split_outer(lg, mid, rs)) return false;
        out = PairPlan{lg, L, mid};
        return true;
    }

    // Transform-count cost of a plan: an unchunked product runs 3 transforms of 2^lg points (2
    // when squaring), a chunked one 1 + 2 * chunks transforms of its shorter length; a transform
    // is priced n * lg.  -1 for a plan that does not apply.  The heuristic picks by this cost
    // (chunked plans must win by 10% to cover their extra launches and Garner reads); the learned
    // model predicts each plan's time relative to it.
    static double plan_cost(uint32_t la, uint32_t lb, bool sq, int id) {
        PairPlan c{};
        if (!plan_for_id(la, lb, sq, id, c)) return -1.0;
        const double tr = c.chunk_len
                              ? 1.0 + 2.0 * (double)((words_of(std::max(la, lb)) + c.chunk_len - 1) / c.chunk_len)
                              : (sq ? 2.0 : 3.0);
        return tr * std::ldexp((double)c.lg, (int)c.lg);
    }

    static PairPlan heuristic_plan(uint32_t la, uint32_t lb, bool sq) {
        PairPlan best{};
        double best_cost = 1e300;
        if (plan_for_id(la, lb, sq, 0, best)) best_cost = plan_cost(la, lb, sq, 0);
        else best = PairPlan{26, kSplitChunk, 12, true};  // beyond one transform
        for (int id = 3; id < kNumPlanIds; ++id) {
            PairPlan c{};
            if (!plan_for_id(la, lb, sq, id, c)) continue;
            const double cost = 1.1 * plan_cost(la, lb, sq, id);
            if (cost < best_cost) {
                best_cost = cost;
                best = c;
            }
        }
        return best;
    }

    // The plan id whose plan the heuristic picks (0 when none of the ids matches it).
    static int heuristic_id(uint32_t la, uint32_t lb, bool sq) {
        const PairPlan h = heuristic_plan(la, lb, sq);
        for (int id = 0; id < kNumPlanIds; ++id) {
            PairPlan q{};
            if (plan_for_id(la, lb, sq, id, q) && q.lg == h.lg && q.chunk_len == h.chunk_len && q.mid_lr == h.mid_lr)
                return id;
        }
        return 0;
    }

    // Learned mode keeps the fallback's plan (table, or heuristic) unless the model predicts a
    // gain over it of more than kModelMargin log2 units (about 2.8 %), and keeps it for launch-
    // bound work, where the fallback's plan is predicted to take under kModelMinBatchUs for the
    // whole batch: there launch costs make plan timings irregular and the measured table is the
    // better judge.  Both were chosen on the training timings (the lowest mean regret over a grid
    // of margins 0.04-0.16 and floors 0-2000 us) and checked on the held-out ones.
    static constexpr float kModelMargin = 0.04f;
    static constexpr double kModelMinBatchUs = 200.0;
    // Largest batch size of the model's training shapes (bigint_amd/planner.py sample_shapes).
    static constexpr uint32_t kModelMaxBatch = 1024;

    // installed planner state
    std::vector<int8_t> table;  // [19][19][13] plan ids by rounded (log2 bits a, log2 bits b, log2 batch)
    int mlp_k = 0;
    float mlp_norm[6] = {0, 0, 0, 1, 1, 1};
    float* d_mlp = nullptr;
    size_t d_mlp_words = 0;
    float* d_scratch = nullptr;
    size_t d_scratch_words = 0;

    static int bucket(double v, int lo, int hi) {
        int b = (int)std::lround(v);
        return std::min(hi, std::max(lo, b)) - lo;
    }
    int table_plan(uint32_t la, uint32_t lb, uint32_t batch) const {
        if (table.empty()) return -1;
        const int ia = bucket(std::log2(28.0 * la), 10, 28), ib = bucket(std::log2(28.0 * lb), 10, 28);
        const int ik = bucket(std::log2((double)std::max(1u, batch)), 0, 12);
        return table[((size_t)ia * 19 + ib) * 13 + ik];
    }

    float* scratch(size_t words) {
        if (words > d_scratch_words) {
            if (d_scratch && hipDeviceSynchronize() != hipSuccess) return nullptr;
            graph_reset();
            if (d_scratch) (void)hipFree(d_scratch);
            d_scratch = nullptr;
            d_scratch_words = 0;
            if (call_malloc((void**)&d_scratch, words * 4) != hipSuccess) return nullptr;
            d_scratch_words = words;
        }
        return d_scratch;
    }

    // Learned choice for a list of shapes: MLP forward + masked argmin on the GPU.
    int mlp_choose_ids(const std::vector<std::array<uint32_t, 3>>& shapes, const std::vector<bool>& sq,
                       const std::vector<uint8_t>& allowed, std::vector<int32_t>& ids) {
        const int B = (int)shapes.size(), K = mlp_k;
        const hipStream_t st = plan_stream;
        ids.assign(B, 0);
        if (B == 0 || K == 0 || !d_mlp) return BI_HIP_EINVAL;
        const size_t words = (size_t)B * 3 + 2 * (size_t)B * K + ((size_t)B * K + 3) / 4 + 4 + 3 * (size_t)B;
        float* w = scratch(words + 64);
        if (!w) return BI_HIP_ENOMEM;
        std::vector<float> x((size_t)B * 3), off((size_t)B * K, 0.f), lbatch((size_t)B);
        std::vector<int32_t> prior((size_t)B);
        for (int i = 0; i < B; ++i) {
            // the fallback's plan: the table's, or the heuristic's where the table defers to it
            int t = table_plan(shapes[i][0], shapes[i][1], shapes[i][2]);
            PairPlan tmp;
            if (t < 0 || t >= K || !plan_for_id(shapes[i][0], shapes[i][1], sq[i], t, tmp))
                t = heuristic_id(shapes[i][0], shapes[i][1], sq[i]);
            prior[i] = t < K ? t : -1;
            // batch sizes beyond the training range (per-product times are flat there: the GPU
            // is full) are presented as the range's largest
            const double f[3] = {std::log2(28.0 * shapes[i][0]), std::log2(28.0 * shapes[i][1]),
                                 std::log2((double)std::min(kModelMaxBatch, std::max(1u, shapes[i][2])))};
            for (int j = 0; j < 3; ++j) x[(size_t)i * 3 + j] = (float)((f[j] - mlp_norm[j]) / mlp_norm[3 + j]);
            lbatch[i] = (float)std::log2((double)std::max(1u, shapes[i][2]));
            // the model predicts log2(time / transform-count cost) per plan
            for (int k = 0; k < K; ++k) {
                const double c = plan_cost(shapes[i][0], shapes[i][1], sq[i], k);
                off[(size_t)i * K + k] = c > 0 ? (float)std::log2(c) : 0.f;
            }
        }
        float* dx = w;
        float* dy = dx + (size_t)B * 3;
        float* doff = dy + (size_t)B * K;
        uint8_t* dallow = reinterpret_cast<uint8_t*>(doff + (size_t)B * K);
        int32_t* dchoice = reinterpret_cast<int32_t*>(doff + (size_t)B * K + ((size_t)B * K + 3) / 4 + 1);
        int32_t* dprior = dchoice + B;
        float* dlbatch = reinterpret_cast<float*>(dprior + B);
        if (hipMemcpyAsync(dx, x.data(), x.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(doff, off.data(), off.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(dallow, allowed.data(), allowed.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(dprior, prior.data(), prior.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(dlbatch, lbatch.data(), lbatch.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess)
            return BI_HIP_EDEVICE;
        if (mlp_forward(dx, B, d_mlp, K, dy, nullptr, nullptr, st) != hipSuccess) return BI_HIP_EDEVICE;
        if (mlp_choose(dy, doff, dallow, dprior, kModelMargin, dlbatch, (float)std::log2(kModelMinBatchUs), B, K,
                       dchoice, st) != hipSuccess)
            return BI_HIP_EDEVICE;
        if (hipMemcpyAsync(ids.data(), dchoice, (size_t)B * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
            return BI_HIP_EDEVICE;
        return hipStreamSynchronize(st) == hipSuccess ? BI_HIP_OK : BI_HIP_EDEVICE;
    }

    // shapes (la, lb, batch) of the pairs; batch = number of pairs sharing the plain length
    static std::vector<std::array<uint32_t, 3>> batch_shapes(const bi_hip_dpair* pairs, size_t count) {
        std::map<int, uint32_t> per_lg;
        std::vector<std::array<uint32_t, 3>> shapes(count);
        for (size_t i = 0; i < count; ++i) {
            const bi_hip_dpair& p = pairs[i];
            const bool sq = p.b == nullptr || (p.b == p.a && p.lb == p.la);
            shapes[i] = {(uint32_t)p.la, (uint32_t)(sq ? p.la : p.lb), 0};
            per_lg[bi_hip_transform_log2(shapes[i][0], shapes[i][1])]++;
        }
        for (size_t i = 0; i < count; ++i) shapes[i][2] = per_lg[bi_hip_transform_log2(shapes[i][0], shapes[i][1])];
        return shapes;
    }

    // Plan ids for shapes under a mode (-1 = heuristic).
    int choose_ids(const std::vector<std::array<uint32_t, 3>>& shapes, const std::vector<bool>& sqv,
                   int mode, std::vector<int32_t>& ids, hipStream_t st) {
        const size_t count = shapes.size();
        ids.assign(count, -1);
        if (mode == BI_HIP_PLAN_LEARNED && mlp_k > 0) {
            // The choice depends on the shape only, and a product-tree level repeats shapes by the
            // thousand: the model runs once per distinct (la, lb, batch, square).
            std::vector<uint32_t> slot(count);
            std::vector<std::array<uint32_t, 3>> ushapes;
            std::vector<bool> usq;
            {
                size_t cap = 16;
                while (cap < 2 * count) cap <<= 1;
                std::vector<uint32_t> table_idx(cap, ~0u);  // open addressing over the distinct shapes
                for (size_t i = 0; i < count; ++i) {
                    const uint64_t h = (((uint64_t)shapes[i][0] * 0x9E3779B97F4A7C15ull) ^
                                        ((uint64_t)shapes[i][1] * 0xC2B2AE3D27D4EB4Full) ^
                                        ((uint64_t)shapes[i][2] << 1) ^ (uint64_t)sqv[i]) * 0x165667B19E3779F9ull;
                    size_t j = (size_t)(h >> 32) & (cap - 1);
                    while (true) {
                        const uint32_t u = table_idx[j];
                        if (u == ~0u) {
                            table_idx[j] = (uint32_t)ushapes.size();
                            slot[i] = (uint32_t)ushapes.size();
                            ushapes.push_back(shapes[i]);
                            usq.push_back(sqv[i]);
                            break;
                        }
                        if (ushapes[u] == shapes[i] && usq[u] == sqv[i]) {
                            slot[i] = u;
                            break;
                        }
                        j = (j + 1) & (cap - 1);
                    }
                }
            }
            const size_t nu = ushapes.size();
            std::vector<uint8_t> allowed(nu * (size_t)mlp_k, 0);
            for (size_t i = 0; i < nu; ++i) {
                PairPlan tmp;
                for (int k = 0; k < mlp_k && k < kNumPlanIds; ++k)
                    allowed[i * mlp_k + k] = plan_for_id(ushapes[i][0], ushapes[i][1], usq[i], k, tmp) ? 1 : 0;
            }
            std::vector<int32_t> uids;
            if (int rc = mlp_choose_ids(ushapes, usq, allowed, uids)) return rc;
            for (size_t i = 0; i < count; ++i) ids[i] = uids[slot[i]];
            return BI_HIP_OK;
        }
        if (mode == BI_HIP_PLAN_TABLE || mode == BI_HIP_PLAN_LEARNED)
            for (size_t i = 0; i < count; ++i) ids[i] = table_plan(shapes[i][0], shapes[i][1], shapes[i][2]);
        return BI_HIP_OK;
    }

    // Per-pair plans for a batch under a plan mode; explicit ids win when given.
    int choose_plans(const bi_hip_dpair* pairs, size_t count, int mode, const int32_t* explicit_ids,
                     std::vector<PairPlan>& out, hipStream_t st) {
        out.resize(count);
        const auto shapes = batch_shapes(pairs, count);
        std::vector<bool> sqv(count);
        for (size_t i = 0; i < count; ++i)
            sqv[i] = pairs[i].b == nullptr || (pairs[i].b == pairs[i].a && pairs[i].lb == pairs[i].la);
        std::vector<int32_t> ids;
        if (explicit_ids) {
            ids.assign(explicit_ids, explicit_ids + count);
        } else {
            const int rc = choose_ids(shapes, sqv, mode, ids, st);
            if (rc) return rc;
        }
        for (size_t i = 0; i < count; ++i) {
            PairPlan pp{};
            if (ids[i] < 0 || !plan_for_id(shapes[i][0], shapes[i][1], sqv[i], ids[i], pp))
                pp = heuristic_plan(shapes[i][0], shapes[i][1], sqv[i]);
            out[i] = pp;
        }
        if (!explicit_ids && count > 1) {
            // Unchunked products of one transform length share a launch set only if they share
            // the block-pass size: give them the size most of them were assigned, so per-pair
            // choices do not split a level into several under-filled launch sets.
            std::map<uint32_t, std::array<uint32_t, 13>> votes;
            for (const PairPlan& pp : out)
                if (!pp.chunk_len) votes[pp.lg][std::min<uint32_t>(pp.mid_lr, 12)]++;
            for (PairPlan& pp : out) {
                if (pp.chunk_len) continue;
                const auto& v = votes[pp.lg];
                pp.mid_lr = (uint32_t)(std::max_element(v.begin(), v.end()) - v.begin());
            }
        }
        return BI_HIP_OK;
    }

    // Batched products on device pointers.  Pairs are grouped by (transform length, block-pass
    // size, kind); one group is one launch sequence with a job per product (per chunk for chunked
    // products).  The shared B transforms of chunked products run first, into a region that
    // lives until their chunk groups have run.
    int run_batch(const bi_hip_dpair* pairs, size_t count, int mode, const int32_t* explicit_ids,
                  hipStream_t st) {
        std::vector<PairPlan> chosen;
        {
            const int rc = choose_plans(pairs, count, mode, explicit_ids, chosen, st);
            if (rc) return rc;
        }
        struct Item {
            uint32_t idx, la, lb;  // la is the operand that gets chunked (the longer one)
            bool sq, swap;
            PairPlan pp;
            uint32_t nchunks;
            size_t bhat;  // word offset of the shared B transform (chunked products)
        };
        using Key = std::tuple<int, uint32_t, uint32_t, int>;  // phase, lg, mid_lr, kind
        std::map<Key, std::vector<uint32_t>> groups;           // -> item indices
        std::vector<Item> items(count);
        std::vector<uint32_t> split;   // products run one by one with both operands chunked
        std::vector<uint32_t> pieces;  // products run one by one in pieces (run_pieces)
        size_t bhat_words = 0;
        for (size_t i = 0; i < count; ++i) {
            const bi_hip_dpair& p = pairs[i];
            const bool sq = p.b == nullptr || (p.b == p.a && p.lb == p.la);
            const size_t lb = sq ? p.la : p.lb;
            if (!p.a || p.la == 0 || lb == 0 || !p.out) return BI_HIP_EINVAL;
            if (p.out_cap < p.la + lb) return BI_HIP_ECAP;
            if (!product_supported(p.la, lb)) return BI_HIP_ETOOBIG;
            Item it{(uint32_t)i, (uint32_t)p.la, (uint32_t)lb, sq, false, chosen[i], 1, 0};
            const uint32_t lg = it.pp.lg, mid = it.pp.mid_lr;
            if (needs_pieces(p.la, lb)) {
                pieces.push_back((uint32_t)i);
            } else if (it.pp.split_both) {
                split.push_back((uint32_t)i);
            } else if (it.pp.chunk_len) {
                it.swap = it.la < it.lb;
                if (it.swap) std::swap(it.la, it.lb);
                it.nchunks = (words_of(it.la) + it.pp.chunk_len - 1) / it.pp.chunk_len;
                it.bhat = bhat_words;
                bhat_words += (size_t)3 << lg;
                groups[{0, lg, mid, kBForward}].push_back((uint32_t)i);
                groups[{1, lg, mid, kChunk}].push_back((uint32_t)i);
            } else {
                groups[{1, lg, mid, sq ? kSquareKind : kPlain}].push_back((uint32_t)i);
            }
            items[i] = it;
        }
        // workspace: [shared B transforms][one group's residues + status words]
        size_t grp_words = 0, desc_need = 0;
        for (auto& [key, idx] : groups) {
            const size_t n = (size_t)1 << std::get<1>(key);
            const int kind = std::get<3>(key);
            size_t words = 0, njobs = 0;
            for (uint32_t i : idx) {
                const Item& it = items[i];
                const size_t nj = kind == kBForward ? 1 : it.nchunks;
                njobs += nj;
                if (kind == kBForward) continue;
                words += 3 * n * (kind == kPlain ? 2 : nj) + carry_tiles((size_t)it.la + it.lb) + 1;
            }
            grp_words = std::max(grp_words, words);
            desc_need += njobs * sizeof(PassJob) + idx.size() * sizeof(GarnerJob) + 512;
        }
        uint32_t* w = workspace(bhat_words + grp_words + 64);
        if (!w) return BI_HIP_ENOMEM;
        if (desc_need > desc_bytes) {
            if (desc && hipDeviceSynchronize() != hipSuccess) return BI_HIP_EDEVICE;
            graph_reset();
            if (desc) (void)hipFree(desc);
            desc = nullptr;
            desc_bytes = 0;
            if (call_malloc((void**)&desc, desc_need) != hipSuccess) return BI_HIP_ENOMEM;
            desc_bytes = desc_need;
        }
        std::vector<uint8_t> hdesc(desc_need);
        struct Launch {
            Plan* pl;
            Kind kind;
            uint32_t njobs, ngjobs, max_out;
            size_t pj_off, gj_off;
            double in_limbs, out_limbs;
            uint32_t status_words;
            uint32_t* status;
        };
        std::vector<Launch> launches;
        size_t off = 0;
        for (auto& [key, idx] : groups) {
            Plan* pl = plan(std::get<1>(key), std::get<2>(key));
            if (!pl) return BI_HIP_ENOMEM;
            const size_t n = (size_t)1 << std::get<1>(key);
            const Kind kind = (Kind)std::get<3>(key);
            Launch L{pl, kind, 0, 0, 0, 0, 0, 0.0, 0.0, 0, nullptr};
            for (uint32_t i : idx) L.njobs += kind == kBForward ? 1 : items[i].nchunks;
            L.ngjobs = kind == kBForward ? 0 : (uint32_t)idx.size();
            L.pj_off = off;
            off += L.njobs * sizeof(PassJob);
            off = (off + 255) & ~(size_t)255;
            L.gj_off = off;
            off += L.ngjobs * sizeof(GarnerJob);
            off = (off + 255) & ~(size_t)255;
            auto* pj = reinterpret_cast<PassJob*>(hdesc.data() + L.pj_off);
            auto* gj = reinterpret_cast<GarnerJob*>(hdesc.data() + L.gj_off);
            uint32_t* res = w + bhat_words;
            size_t res_words = 0;
            if (kind != kBForward)
                for (uint32_t i : idx) res_words += 3 * n * (kind == kPlain ? 2 : items[i].nchunks);
            uint32_t* status = res + res_words;
            uint32_t* st_cur = status;
            size_t jp = 0;
            for (size_t j = 0; j < idx.size(); ++j) {
                const Item& it = items[idx[j]];
                const bi_hip_dpair& p = pairs[it.idx];
                const uint32_t* A = it.swap ? p.b : p.a;
                const uint32_t* B = it.sq ? p.a : (it.swap ? p.a : p.b);
                if (kind == kBForward) {
                    pj[jp++] = PassJob{B, nullptr, it.lb, 0, w + it.bhat, nullptr, 0u, ~0u};
                    L.in_limbs += (double)it.lb;
                    continue;
                }
                GarnerJob g{};
                g.y = res;
                if (kind == kChunk) {
                    const uint32_t Lc = it.pp.chunk_len;
                    for (uint32_t c = 0; c < it.nchunks; ++c) {
                        pj[jp++] = PassJob{A, nullptr, it.la, it.lb, res, w + it.bhat, c * Lc, Lc};
                        res += 3 * n;
                    }
                    g.nchunks = it.nchunks;
                    g.chunk_len = Lc;
                    g.log_chunk = (Lc & (Lc - 1)) == 0 ? ceil_log2(Lc) : 0xFFu;
                    g.chunk_stride = 3 * n;
                } else {
                    uint32_t* ra = res;
                    res += 3 * n;
                    uint32_t* rb = nullptr;
                    if (!it.sq) {
                        rb = res;
                        res += 3 * n;
                    }
                    pj[jp++] = PassJob{A, B, it.la, it.lb, ra, rb, 0u, ~0u};
                    g.nchunks = 1;
                    if ((uint64_t)words_of(it.la) + words_of(it.lb) - 1 > n) {
                        g.wrap_a = A;
                        g.wrap_b = B;
                        g.wrap_la = it.la;
                        g.wrap_lb = it.lb;
                    }
                }
                const uint32_t out_len = it.la + it.lb;
                const uint32_t nt = carry_tiles(out_len);
                g.out = p.out;
                g.n = (uint32_t)n;
                g.out_len = out_len;
                for (int I = 0; I < 3; ++I) g.kscale[I] = pl->kscale[I];
                g.status = st_cur;
                g.fix = st_cur + nt;
                st_cur += nt + 1;
                gj[j] = g;
                L.max_out = std::max(L.max_out, out_len);
                L.in_limbs += (double)it.la + (kind == kPlain ? (double)it.lb : 0.0);
                L.out_limbs += (double)out_len;
            }
            L.status_words = (uint32_t)(st_cur - status);
            L.status = status;
            launches.push_back(L);
        }
        // the descriptors of earlier calls may still be read, on this stream or another: the new
        // ones are copied in after that work, in stream order
        if (int rc = upload_desc(hdesc.data(), off, st)) return rc;
        auto launch_all = [&]() -> int {
            for (auto& L : launches) {
                const auto* pj = reinterpret_cast<const PassJob*>(desc + L.pj_off);
                const auto* gj = reinterpret_cast<const GarnerJob*>(desc + L.gj_off);
                const auto* hp = reinterpret_cast<const PassJob*>(hdesc.data() + L.pj_off);
                const GarnerJob g0 = L.ngjobs ? reinterpret_cast<const GarnerJob*>(hdesc.data() + L.gj_off)[0]
                                              : GarnerJob{};
                const int rc = launch_products(L.pl, L.kind, L.njobs, hp[0], L.njobs == 1 ? nullptr : pj,
                                               L.ngjobs, g0, L.ngjobs == 1 ? nullptr : gj, L.max_out,
                                               L.in_limbs, L.out_limbs, L.status, L.status_words, st);
                if (rc) return rc;
            }
            return BI_HIP_OK;
        };
        // A level's launch sets go to the GPU as one captured graph (kernels only: the status
        // words are cleared by a fill kernel, not a memset node).  A graph of the same shape as
        // the previous one is updated in place instead of instantiated again.
        const bool use_graph = graphs && prof_mask == 0 && launches.size() > 1;
        const int rc = use_graph ? launch_as_graph(launch_all, st) : launch_all();
        if (rc) return rc;
        if (int r2 = used_shared(st)) return r2;
        for (uint32_t i : split) {
            const bi_hip_dpair& p = pairs[i];
            const bool sq = p.b == nullptr || (p.b == p.a && p.lb == p.la);
            const int rc = run_split_both(p.a, (uint32_t)p.la, sq ? p.a : p.b, (uint32_t)(sq ? p.la : p.lb),
                                          sq, p.out, st);
            if (rc) return rc;
        }
        for (uint32_t i : pieces) {
            const bi_hip_dpair& p = pairs[i];
            const bool sq = p.b == nullptr || (p.b == p.a && p.lb == p.la);
            const int rc = run_pieces(p.a, (uint32_t)p.la, sq ? p.a : p.b, (uint32_t)(sq ? p.la : p.lb), p.out, st);
            if (rc) return rc;
        }
        return BI_HIP_OK;
    }

    // One product whose shorter operand has more than 2^26 words (see needs_pieces): piece i of
    // the shorter operand B (limbs [i K, i K + K), K = kPieceLimbs) times A is computed by run()
    // and added into out at limb offset i K, where the partial sum of the pieces before it has
    // exactly la significant limbs; piece 0's product is written to out directly.  The partial
    // products and the additions' scratch live in one allocation freed before returning, so this
    // tier synchronises st.
    int run_pieces(const uint32_t* A, uint32_t la, const uint32_t* B, uint32_t lb, uint32_t* out, hipStream_t st) {
        if (la < lb) {
            std::swap(A, B);
            std::swap(la, lb);
        }
        const uint32_t np = (uint32_t)(((uint64_t)lb + kPieceLimbs - 1) / kPieceLimbs);
        const size_t tmp_limbs = (size_t)la + kPieceLimbs;
        const size_t ntiles = (tmp_limbs + kAddTileLimbs - 1) / kAddTileLimbs;
        const size_t jobs_off = (tmp_limbs * 4 + 255) & ~(size_t)255;
        const size_t words_off = jobs_off + (((size_t)np * sizeof(AddJob) + 255) & ~(size_t)255);
        uint8_t* buf = nullptr;
        if (call_malloc((void**)&buf, words_off + (2 * ntiles + 2) * 4) != hipSuccess) return BI_HIP_ENOMEM;
        uint32_t* tmp = reinterpret_cast<uint32_t*>(buf);
        uint32_t* maps = reinterpret_cast<uint32_t*>(buf + words_off);
        std::vector<AddJob> jobs(np);
        for (uint32_t i = 1; i < np; ++i) {
            const size_t off = (size_t)i * kPieceLimbs, lp = std::min<size_t>(kPieceLimbs, lb - off);
            const size_t width = (size_t)la + lp;  // the sum is below 2^(28 (la + lp)) at offset off
            jobs[i] = AddJob{out + off, la, tmp, width, out + off, width, maps, maps + ntiles, maps + 2 * ntiles,
                             0u, (uint32_t)((width + kAddTileLimbs - 1) / kAddTileLimbs)};
        }
        auto* d_jobs = reinterpret_cast<AddJob*>(buf + jobs_off);
        int rc = hipMemcpyAsync(d_jobs, jobs.data(), np * sizeof(AddJob), hipMemcpyHostToDevice, st) == hipSuccess
                     ? BI_HIP_OK
                     : BI_HIP_EDEVICE;
        for (uint32_t i = 0; i < np && rc == BI_HIP_OK; ++i) {
            const size_t off = (size_t)i * kPieceLimbs;
            const uint32_t lp = (uint32_t)std::min<size_t>(kPieceLimbs, lb - off);
            rc = run(A, la, B + off, lp, false, i == 0 ? out : tmp, st);
            // the add kernels read the partial product written by run()'s last kernel on st
            if (rc == BI_HIP_OK && i > 0 && launch_add_batch(d_jobs + i, 1, jobs[i].ntiles, st) != hipSuccess)
                rc = BI_HIP_EDEVICE;
        }
        // the buffer is read by kernels queued on st: they finish before it is freed
        if (hipStreamSynchronize(st) != hipSuccess && rc == BI_HIP_OK) rc = BI_HIP_EDEVICE;
        (void)hipFree(buf);
        return rc;
    }

    // One product with both operands cut into kSplitChunk-word chunks (see PairPlan): forward
    // transforms of every chunk (the transforms of a square's B are A's), diagonal sums in the
    // transform domain, one inverse transform per diagonal, and Garner summing the diagonals,
    // which overlap like the chunk products of a chunked plan.
    int run_split_both(const uint32_t* A, uint32_t la, const uint32_t* B, uint32_t lb, bool sq,
                       uint32_t* out, hipStream_t st) {
        const uint32_t lg = 26, L = kSplitChunk;
        const size_t n = (size_t)1 << lg;
        const uint32_t ca = (words_of(la) + L - 1) / L, cb = sq ? ca : (words_of(lb) + L - 1) / L;
        const uint32_t nd = ca + cb - 1;
        Plan* pl = plan(lg, 12);
        if (!pl) return BI_HIP_ENOMEM;
        const uint32_t out_len = la + lb;
        const uint32_t nt = carry_tiles(out_len);
        const size_t nspec = ca + (sq ? 0 : cb);
        uint32_t* w = workspace((nspec + nd) * 3 * n + nt + 64);
        if (!w) return BI_HIP_ENOMEM;
        uint32_t* spec_a = w;
        uint32_t* spec_b = sq ? spec_a : w + (size_t)ca * 3 * n;
        uint32_t* diag = w + nspec * 3 * n;
        uint32_t* status = diag + (size_t)nd * 3 * n;
        std::vector<PassJob> fwd, inv(nd);
        for (uint32_t c = 0; c < ca; ++c)
            fwd.push_back(PassJob{A, nullptr, la, 0, spec_a + (size_t)c * 3 * n, nullptr, c * L, L});
        if (!sq)
            for (uint32_t c = 0; c < cb; ++c)
                fwd.push_back(PassJob{B, nullptr, lb, 0, spec_b + (size_t)c * 3 * n, nullptr, c * L, L});
        for (uint32_t s = 0; s < nd; ++s)
            inv[s] = PassJob{nullptr, nullptr, 0, 0, diag + (size_t)s * 3 * n, nullptr, 0u, ~0u};
        const size_t fwd_bytes = fwd.size() * sizeof(PassJob), inv_off = (fwd_bytes + 255) & ~(size_t)255;
        const size_t need = inv_off + inv.size() * sizeof(PassJob);
        if (need > desc_bytes) {
            if (desc && hipDeviceSynchronize() != hipSuccess) return BI_HIP_EDEVICE;
            graph_reset();
            if (desc) (void)hipFree(desc);
            desc = nullptr;
            desc_bytes = 0;
            if (call_malloc((void**)&desc, need) != hipSuccess) return BI_HIP_ENOMEM;
            desc_bytes = need;
        }
        if (wait_shared() || hipStreamSynchronize(st) != hipSuccess ||
            hipMemcpy(desc, fwd.data(), fwd_bytes, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(desc + inv_off, inv.data(), inv.size() * sizeof(PassJob), hipMemcpyHostToDevice) != hipSuccess)
            return BI_HIP_EDEVICE;
        const auto* d_fwd = reinterpret_cast<const PassJob*>(desc);
        const auto* d_inv = reinterpret_cast<const PassJob*>(desc + inv_off);
        GarnerJob g{};
        int rc = launch_products(pl, kBForward, (uint32_t)fwd.size(), fwd[0], fwd.size() == 1 ? nullptr : d_fwd,
                                 0, g, nullptr, 0, (double)la + (sq ? 0.0 : (double)lb), 0.0, nullptr, 0, st);
        if (rc) return rc;
        if (timed(BI_HIP_K_MID, 12.0 * n * (2.0 * ca * cb + nd), st, [&] {
                return launch_diag_accumulate(spec_a, spec_b, ca, cb, (uint32_t)n, diag, st);
            }) != hipSuccess)
            return BI_HIP_EDEVICE;
        g.y = diag;
        g.out = out;
        g.n = (uint32_t)n;
        g.out_len = out_len;
        for (int I = 0; I < 3; ++I) g.kscale[I] = pl->kscale[I];
        g.status = status;
        g.fix = status + nt;
        g.nchunks = nd;
        g.chunk_len = L;
        g.log_chunk = kSplitLog;
        g.chunk_stride = 3 * n;
        rc = launch_products(pl, kInverse, nd, inv[0], nd == 1 ? nullptr : d_inv, 1, g, nullptr, out_len,
                             0.0, (double)out_len, status, nt + 1, st);
        if (rc) return rc;
        return used_shared(st);
    }
};

Engine& engine() {
    static Engine e;
    return e;
}

}  // namespace
}  // namespace bi

using bi::engine;

extern "C" {

int bi_hip_init(int device) {
    std::lock_guard<std::mutex> lk(engine().mtx);
    return engine().init(device);
}

int bi_hip_transform_log2(size_t la, size_t lb) {
    if (la == 0 || lb == 0 || la > 0xFFFFFFFFull || lb > 0xFFFFFFFFull) return -1;
    // The transforms run over 32-bit words of the operands (8 limbs of 28 bits = 7 words).  The
    // product has wa + wb - 1 coefficients; the top one may wrap onto coefficient 0 (Garner
    // separates them again) as long as both operands fit the transform.
    const size_t wa = (la * 7 + 7) / 8, wb = (lb * 7 + 7) / 8;
    const size_t need = std::max(wa + wb - 2, std::max(wa, wb));
    if (need > ((size_t)1 << bi::kMaxLog)) return -1;
    int lg = 4;  // at least one 16-point tile
    while (((size_t)1 << lg) < need) ++lg;
    return lg;
}

int bi_hip_mul_device(const uint32_t* d_a, size_t la, const uint32_t* d_b, size_t lb,
                      int is_square, uint32_t* d_out, size_t out_cap, void* stream) {
    if (!d_a || la == 0 || (!is_square && (!d_b || lb == 0)) || !d_out) return BI_HIP_EINVAL;
    if (is_square) lb = la;
    if (out_cap < la + lb) return BI_HIP_ECAP;
    // lengths are checked as size_t before run() takes them as 32-bit counts
    if (!bi::Engine::product_supported(la, lb)) return BI_HIP_ETOOBIG;
    std::lock_guard<std::mutex> lk(engine().mtx);
    int rc = engine().init(0);
    if (rc) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : engine().stream;
    return engine().run(d_a, (uint32_t)la, is_square ? d_a : d_b, (uint32_t)lb, is_square != 0,
                        d_out, st);
}

int bi_hip_mul_device_noalloc(const uint32_t* d_a, size_t la, const uint32_t* d_b, size_t lb,
                              int is_square, uint32_t* d_out, size_t out_cap, void* stream) {
    if (!d_a || la == 0 || (!is_square && (!d_b || lb == 0)) || !d_out) return BI_HIP_EINVAL;
    if (is_square) lb = la;
    if (out_cap < la + lb) return BI_HIP_ECAP;
    std::lock_guard<std::mutex> lk(engine().mtx);
    bi::Engine& e = engine();
    if (!e.ready || !e.noalloc_scratch) return BI_HIP_ENOINIT;
    const int lg = bi_hip_transform_log2(la, lb);
    if (lg < 0 || lg > (int)kNoAllocMaxLog) return BI_HIP_ESCRATCH;
    if (bi::Engine::single_words((uint32_t)lg, (uint32_t)la, (uint32_t)lb, is_square != 0) > kNoAllocWords)
        return BI_HIP_ESCRATCH;
    const bi::Plan* pl = e.find_plan((uint32_t)lg);
    if (!pl) return BI_HIP_ENOINIT;
    hipStream_t st = stream ? (hipStream_t)stream : e.stream;
    e.suppress_prof = true;  // profiling would create events
    const int rc = e.launch_single(pl, d_a, (uint32_t)la, is_square ? d_a : d_b, (uint32_t)lb, is_square != 0,
                                   d_out, e.noalloc_scratch, st);
    e.suppress_prof = false;
    return rc;
}

namespace {

// a * b for host limb arrays; limbs [la + lb - keep, la + lb) of the product go to out.
int mul_host(const uint32_t* a, size_t la, const uint32_t* b, size_t lb, int is_square,
             size_t keep, uint32_t* out) {
    std::lock_guard<std::mutex> lk(engine().mtx);
    bi::Engine& e = engine();
    int rc = e.init(0);
    if (rc) return rc;
    const size_t words = la + (is_square ? 0 : lb) + la + lb;
    if (words > e.io_words) {
        if (e.io && hipDeviceSynchronize() != hipSuccess) return BI_HIP_EDEVICE;
        e.graph_reset();
        if (e.io) (void)hipFree(e.io);
        e.io = nullptr;
        e.io_words = 0;
        if (bi::call_malloc((void**)&e.io, words * 4) != hipSuccess) return BI_HIP_ENOMEM;
        e.io_words = words;
    }
    uint32_t* da = e.io;
    uint32_t* db = is_square ? da : da + la;
    uint32_t* dout = da + la + (is_square ? 0 : lb);
    std::lock_guard<std::mutex> plk(e.pipe_mtx);
    if (e.pipe.h2d(da, a, la * 4, e.stream) != hipSuccess) return BI_HIP_EDEVICE;
    if (!is_square && e.pipe.h2d(db, b, lb * 4, e.stream) != hipSuccess) return BI_HIP_EDEVICE;
    rc = e.run(da, (uint32_t)la, db, (uint32_t)lb, is_square != 0, dout, e.stream);
    if (rc) return rc;
    if (e.pipe.d2h(out, dout + (la + lb - keep), keep * 4, e.stream) != hipSuccess) return BI_HIP_EDEVICE;
    return BI_HIP_OK;
}

int check_host_product(const uint32_t* a, size_t la, const uint32_t*& b, size_t& lb, int is_square) {
    if (!a || la == 0 || (!is_square && (!b || lb == 0))) return BI_HIP_EINVAL;
    if (is_square) {
        b = a;
        lb = la;
    }
    if (a[la - 1] == 0 || b[lb - 1] == 0) return BI_HIP_EINVAL;
    if (!bi::Engine::product_supported(la, lb)) return BI_HIP_ETOOBIG;
    return BI_HIP_OK;
}

}  // namespace

int bi_hip_mul(const uint32_t* a, size_t la, const uint32_t* b, size_t lb, int is_square,
               uint32_t* out, size_t out_cap, size_t* out_len) {
    if (!out || !out_len) return BI_HIP_EINVAL;
    if (int rc = check_host_product(a, la, b, lb, is_square)) return rc;
    if (out_cap < la + lb) return BI_HIP_ECAP;
    if (int rc = mul_host(a, la, b, lb, is_square, la + lb, out)) return rc;
    size_t n = la + lb;
    while (n > 1 && out[n - 1] == 0) --n;
    *out_len = n;
    return BI_HIP_OK;
}

int bi_hip_mul_top(const uint32_t* a, size_t la, const uint32_t* b, size_t lb, int is_square,
                   size_t keep, uint32_t* out, size_t* out_len) {
    if (!out || !out_len || keep == 0) return BI_HIP_EINVAL;
    if (int rc = check_host_product(a, la, b, lb, is_square)) return rc;
    const size_t k = std::min(keep, la + lb);
    if (int rc = mul_host(a, la, b, lb, is_square, k, out)) return rc;
    // operands with nonzero top limbs: the product has la + lb or la + lb - 1 limbs
    *out_len = la + lb - (out[k - 1] == 0 && la + lb > 1 ? 1 : 0);
    return BI_HIP_OK;
}

int bi_hip_mul_batch_device(const bi_hip_dpair* pairs, size_t count, int plan_mode,
                            void* stream) {
    if (!pairs && count) return BI_HIP_EINVAL;
    if (count == 0) return BI_HIP_OK;
    std::lock_guard<std::mutex> lk(engine().mtx);
    bi::Engine& e = engine();
    int rc = e.init(0);
    if (rc) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : e.stream;
    if (plan_mode == BI_HIP_PLAN_FIXED) {
        for (size_t i = 0; i < count; ++i) {
            const bi_hip_dpair& p = pairs[i];
            const bool sq = p.b == nullptr || (p.b == p.a && p.lb == p.la);
            const size_t lb = sq ? p.la : p.lb;
            if (!p.a || !p.out || p.la == 0 || lb == 0) return BI_HIP_EINVAL;
            if (p.out_cap < p.la + lb) return BI_HIP_ECAP;
            // lengths are checked as size_t before run() takes them as 32-bit counts
            if (!bi::Engine::product_supported(p.la, lb)) return BI_HIP_ETOOBIG;
            rc = e.run(p.a, (uint32_t)p.la, sq ? p.a : p.b, (uint32_t)lb, sq, p.out, st, false);
            if (rc) return rc;
        }
        return BI_HIP_OK;
    }
    return e.run_batch(pairs, count, plan_mode, nullptr, st);
}

int bi_hip_mul_batch(bi_hip_pair* pairs, size_t count, int plan_mode) {
    if (!pairs && count) return BI_HIP_EINVAL;
    if (count == 0) return BI_HIP_OK;
    // validate, stage every operand in one device buffer, run the device batch, copy back
    size_t in_words = 0, out_words = 0;
    for (size_t i = 0; i < count; ++i) {
        bi_hip_pair& p = pairs[i];
        const bool sq = p.a == p.b && p.la == p.lb;
        if (!p.a || p.la == 0 || (!sq && (!p.b || p.lb == 0)) || !p.out) {
            p.status = BI_HIP_EINVAL;
            return BI_HIP_EINVAL;
        }
        if (p.a[p.la - 1] == 0 || p.b[p.lb - 1] == 0) {
            p.status = BI_HIP_EINVAL;
            return BI_HIP_EINVAL;
        }
        if (p.out_cap < p.la + p.lb) {
            p.status = BI_HIP_ECAP;
            return BI_HIP_ECAP;
        }
        if (!bi::Engine::product_supported(p.la, p.lb)) {
            p.status = BI_HIP_ETOOBIG;
            return BI_HIP_ETOOBIG;
        }
        in_words += p.la + (sq ? 0 : p.lb);
        out_words += p.la + p.lb;
    }
    bi::Engine& e = engine();
    {
        std::lock_guard<std::mutex> lk(e.mtx);
        int rc = e.init(0);
        if (rc) return rc;
    }
    // one staging call at a time: the device staging buffer is the engine's, grow-only
    std::lock_guard<std::mutex> stage_lk(e.stage_mtx);
    if (in_words + out_words > e.stage_words) {
        if (e.stage && hipDeviceSynchronize() != hipSuccess) return BI_HIP_EDEVICE;
        e.graph_reset();
        if (e.stage) (void)hipFree(e.stage);
        e.stage = nullptr;
        e.stage_words = 0;
        if (bi::call_malloc((void**)&e.stage, (in_words + out_words) * 4) != hipSuccess) return BI_HIP_ENOMEM;
        e.stage_words = in_words + out_words;
    }
    uint32_t* dbuf = e.stage;
    // operands are gathered into the staging buffer and products scattered out of it through the
    // pinned pipe, with no host-side staging copy
    bi::HostSegs hin, hout;
    std::vector<bi_hip_dpair> dp(count);
    size_t io = 0, oo = in_words;
    for (size_t i = 0; i < count; ++i) {
        bi_hip_pair& p = pairs[i];
        const bool sq = p.a == p.b && p.la == p.lb;
        hin.add(p.a, p.la * 4);
        dp[i].a = dbuf + io;
        dp[i].la = p.la;
        io += p.la;
        if (sq) {
            dp[i].b = dp[i].a;
        } else {
            hin.add(p.b, p.lb * 4);
            dp[i].b = dbuf + io;
            io += p.lb;
        }
        dp[i].lb = p.lb;
        dp[i].out = dbuf + oo;
        dp[i].out_cap = p.la + p.lb;
        oo += p.la + p.lb;
        hout.add(p.out, (p.la + p.lb) * 4);
    }
    int rc = BI_HIP_OK;
    {
        std::lock_guard<std::mutex> plk(e.pipe_mtx);
        if (e.pipe.h2d(dbuf, hin, e.stream) != hipSuccess || hipStreamSynchronize(e.stream) != hipSuccess)
            rc = BI_HIP_EDEVICE;
    }
    if (!rc) rc = bi_hip_mul_batch_device(dp.data(), count, plan_mode, nullptr);
    if (!rc) {
        std::lock_guard<std::mutex> plk(e.pipe_mtx);
        if (e.pipe.d2h(hout, dbuf + in_words, e.stream) != hipSuccess) rc = BI_HIP_EDEVICE;
    }
    for (size_t i = 0; i < count; ++i) {
        bi_hip_pair& p = pairs[i];
        p.status = rc;
        if (rc) continue;
        size_t n = p.la + p.lb;
        while (n > 1 && p.out[n - 1] == 0) --n;
        p.out_len = n;
    }
    return rc;
}

int bi_hip_set_graphs(int enable) {
    std::lock_guard<std::mutex> lk(engine().mtx);
    engine().graphs = enable != 0;
    return BI_HIP_OK;
}

int bi_hip_plan_count(void) { return bi::Engine::kNumPlanIds; }

double bi_hip_plan_cost(size_t la, size_t lb, int is_square, int plan_id) {
    if (la == 0 || lb == 0 || la > 0xFFFFFFFFull || lb > 0xFFFFFFFFull) return -1.0;
    return bi::Engine::plan_cost((uint32_t)la, (uint32_t)lb, is_square != 0, plan_id);
}

int bi_hip_plan_allowed(size_t la, size_t lb, int is_square, int plan_id) {
    if (la == 0 || lb == 0 || la > 0xFFFFFFFFull || lb > 0xFFFFFFFFull) return 0;
    bi::Engine::PairPlan p{};
    return bi::Engine::plan_for_id((uint32_t)la, (uint32_t)lb, is_square != 0, plan_id, p) ? 1 : 0;
}

int bi_hip_mul_batch_device_planned(const bi_hip_dpair* pairs, size_t count,
                                    const int32_t* plan_ids, void* stream) {
    if ((!pairs || !plan_ids) && count) return BI_HIP_EINVAL;
    if (count == 0) return BI_HIP_OK;
    std::lock_guard<std::mutex> lk(engine().mtx);
    bi::Engine& e = engine();
    int rc = e.init(0);
    if (rc) return rc;
    return e.run_batch(pairs, count, BI_HIP_PLAN_TABLE, plan_ids, stream ? (hipStream_t)stream : e.stream);
}

int bi_hip_choose_plans(const uint32_t* shapes, const int* is_square, size_t count, int mode,
                        int32_t* ids) {
    if ((!shapes || !ids) && count) return BI_HIP_EINVAL;
    std::lock_guard<std::mutex> lk(engine().mtx);
    bi::Engine& e = engine();
    int rc = e.init(0);
    if (rc) return rc;
    std::vector<std::array<uint32_t, 3>> sh(count);
    std::vector<bool> sqv(count);
    for (size_t i = 0; i < count; ++i) {
        sh[i] = {shapes[3 * i], shapes[3 * i + 1], shapes[3 * i + 2]};
        sqv[i] = is_square && is_square[i];
    }
    std::vector<int32_t> out;
    rc = e.choose_ids(sh, sqv, mode, out, e.stream);
    if (rc) return rc;
    for (size_t i = 0; i < count; ++i) {
        bi::Engine::PairPlan p{};
        if (out[i] < 0 || !bi::Engine::plan_for_id(sh[i][0], sh[i][1], sqv[i], out[i], p)) {
            // report which id the heuristic corresponds to
            out[i] = bi::Engine::heuristic_id(sh[i][0], sh[i][1], sqv[i]);
        }
        ids[i] = out[i];
    }
    return BI_HIP_OK;
}

int bi_hip_set_planner(const int8_t* table, const float* mlp_params, int K, const float* norm6) {
    std::lock_guard<std::mutex> lk(engine().mtx);
    bi::Engine& e = engine();
    int rc = e.init(0);
    if (rc) return rc;
    if (table) e.table.assign(table, table + 19 * 19 * 13);
    if (mlp_params) {
        if (K <= 0 || K > bi::kMaxPlans || !norm6) return BI_HIP_EINVAL;
        const size_t np = (size_t)bi::mlp_param_count(K);
        if (np > e.d_mlp_words) {
            if (e.d_mlp && hipDeviceSynchronize() != hipSuccess) return BI_HIP_EDEVICE;
            e.graph_reset();
            if (e.d_mlp) (void)hipFree(e.d_mlp);
            e.d_mlp = nullptr;
            e.d_mlp_words = 0;
            if (hipMalloc((void**)&e.d_mlp, np * 4) != hipSuccess) return BI_HIP_ENOMEM;
            e.d_mlp_words = np;
        }
        if (hipMemcpy(e.d_mlp, mlp_params, np * 4, hipMemcpyHostToDevice) != hipSuccess) return BI_HIP_EDEVICE;
        e.mlp_k = K;
        for (int i = 0; i < 6; ++i) e.mlp_norm[i] = norm6[i];
    }
    return BI_HIP_OK;
}

int bi_hip_load_planner(const char* path) {
    if (!path) return BI_HIP_EINVAL;
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return BI_HIP_EINVAL;
    char magic[4];
    int32_t K = 0, np = 0;
    std::vector<int8_t> table(19 * 19 * 13);
    std::vector<float> params, norm(6);
    bool ok = std::fread(magic, 1, 4, f) == 4 && std::memcmp(magic, "BIPL", 4) == 0 &&
              std::fread(&K, 4, 1, f) == 1 && K > 0 && K <= bi::kMaxPlans &&
              std::fread(table.data(), 1, table.size(), f) == table.size() &&
              std::fread(&np, 4, 1, f) == 1 && np == bi::mlp_param_count(K);
    if (ok) {
        params.resize((size_t)np);
        ok = std::fread(params.data(), 4, params.size(), f) == params.size() &&
             std::fread(norm.data(), 4, 6, f) == 6;
    }
    std::fclose(f);
    if (!ok) return BI_HIP_EINVAL;
    for (int8_t t : table)  // plan ids, or -1: the built-in heuristic
        if (t < -1 || t >= K) return BI_HIP_EINVAL;
    return bi_hip_set_planner(table.data(), params.data(), K, norm.data());
}

int bi_hip_mlp_param_count(int K) { return (K > 0 && K <= bi::kMaxPlans) ? bi::mlp_param_count(K) : -1; }

int bi_hip_mlp_forward(const float* d_x, int B, const float* d_params, int K, float* d_y, void* stream) {
    if (K <= 0 || K > bi::kMaxPlans || B < 0 || !d_x || !d_params || !d_y) return BI_HIP_EINVAL;
    return bi::mlp_forward(d_x, B, d_params, K, d_y, nullptr, nullptr, (hipStream_t)stream) == hipSuccess
               ? BI_HIP_OK : BI_HIP_EDEVICE;
}

int bi_hip_mlp_grad(const float* d_x, const float* d_t, const float* d_mask, int B,
                    const float* d_params, int K, float* d_grad, void* stream) {
    if (K <= 0 || K > bi::kMaxPlans || B <= 0 || !d_x || !d_t || !d_mask || !d_params || !d_grad)
        return BI_HIP_EINVAL;
    std::lock_guard<std::mutex> lk(engine().mtx);
    bi::Engine& e = engine();
    int rc = e.init(0);
    if (rc) return rc;
    const int NP = bi::mlp_param_count(K);
    const size_t nb = ((size_t)B + 63) / 64;
    const size_t words = (size_t)B * K + 2 * (size_t)B * 64 + nb * NP + nb + 64;
    float* w = e.scratch(words);
    if (!w) return BI_HIP_ENOMEM;
    float* y = w;
    float* h1 = y + (size_t)B * K;
    float* h2 = h1 + (size_t)B * 64;
    float* part = h2 + (size_t)B * 64;
    float* part_loss = part + nb * NP;
    // count of measured entries, for the mean: computed on the host from the mask
    std::vector<float> hm((size_t)B * K);
    if (hipMemcpy(hm.data(), d_mask, hm.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return BI_HIP_EDEVICE;
    double cnt = 0;
    for (float m : hm) cnt += m;
    const float inv = cnt > 0 ? (float)(1.0 / cnt) : 0.f;
    hipStream_t st = (hipStream_t)stream;
    if (bi::mlp_grad(d_x, d_t, d_mask, B, d_params, K, inv, y, h1, h2, part, part_loss, d_grad, st) != hipSuccess)
        return BI_HIP_EDEVICE;
    return hipStreamSynchronize(st) == hipSuccess ? BI_HIP_OK : BI_HIP_EDEVICE;
}

int bi_hip_mlp_adam(float* d_params, const float* d_grad, float* d_m, float* d_v, int nparams,
                    float lr, float beta1, float beta2, float eps, int step, void* stream) {
    if (!d_params || !d_grad || !d_m || !d_v || nparams <= 0 || step <= 0) return BI_HIP_EINVAL;
    return bi::mlp_adam(d_params, d_grad, d_m, d_v, nparams, lr, beta1, beta2, eps, step,
                        (hipStream_t)stream) == hipSuccess ? BI_HIP_OK : BI_HIP_EDEVICE;
}

int bi_hip_profile_kinds(uint32_t mask) {
    mask &= (1u << BI_HIP_K_COUNT) - 1u;
    std::lock_guard<std::mutex> lk(engine().mtx);
    bi::Engine& e = engine();
    if (mask) {
        if (e.ready && hipStreamSynchronize(e.stream) != hipSuccess) return BI_HIP_EDEVICE;
        e.prof_clear();
    }
    e.prof_mask = mask;
    return BI_HIP_OK;
}

int bi_hip_profile(int enable) { return bi_hip_profile_kinds(enable ? ~0u : 0u); }

int bi_hip_profile_read(bi_hip_kstat* stats, size_t cap) {
    if (!stats || cap < BI_HIP_K_COUNT) return BI_HIP_EINVAL;
    std::lock_guard<std::mutex> lk(engine().mtx);
    bi::Engine& e = engine();
    for (size_t i = 0; i < BI_HIP_K_COUNT; ++i) stats[i] = bi_hip_kstat{0.0, 0.0, 0};
    for (auto& r : e.prof) {
        if (hipEventSynchronize(r.e1) != hipSuccess) return BI_HIP_EDEVICE;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.e0, r.e1) != hipSuccess) return BI_HIP_EDEVICE;
        stats[r.kind].ms += ms;
        stats[r.kind].bytes += r.bytes;
        stats[r.kind].launches += 1;
    }
    return BI_HIP_OK;
}

int bi_hip_add_batch_device(bi_hip_dadd* jobs, size_t count, void* stream) {
    if (!jobs && count) return BI_HIP_EINVAL;
    if (count == 0) return BI_HIP_OK;
    if (count > 0xFFFFu) return BI_HIP_EINVAL;  // one grid row per job
    std::vector<bi::AddJob> hj(count);
    size_t tiles = 0;
    uint32_t max_tiles = 0;
    for (size_t i = 0; i < count; ++i) {
        const bi_hip_dadd& d = jobs[i];
        if (!d.x || !d.y || !d.out || d.lx == 0 || d.ly == 0) return BI_HIP_EINVAL;
        const size_t width = std::max(d.lx, d.ly) + (d.subtract ? 0 : 1);
        if (d.out_cap < width) return BI_HIP_ECAP;
        const size_t nt = (width + bi::kAddTileLimbs - 1) / bi::kAddTileLimbs;
        if (nt > 0xFFFFFFFFull) return BI_HIP_ETOOBIG;
        hj[i] = bi::AddJob{d.x, d.lx, d.y, d.ly, d.out, width, nullptr, nullptr, nullptr,
                           d.subtract ? 1u : 0u, (uint32_t)nt};
        tiles += nt;
        max_tiles = std::max(max_tiles, (uint32_t)nt);
    }
    if (max_tiles > 0x7FFFFFFFu) return BI_HIP_ETOOBIG;
    bi::Engine& e = engine();
    {
        std::lock_guard<std::mutex> lk(e.mtx);
        if (int rc = e.init(0)) return rc;
    }
    std::lock_guard<std::mutex> lk(e.add_mtx);
    hipStream_t st = stream ? (hipStream_t)stream : e.stream;
    // scratch: descriptors, then maps and carries (one word per tile each), then 2 flags per job
    const size_t desc_bytes = (count * sizeof(bi::AddJob) + 255) & ~size_t(255);
    const size_t need = desc_bytes + (2 * tiles + 2 * count) * 4;
    if (need > e.add_bytes) {
        if (e.add_scratch && hipStreamSynchronize(st) != hipSuccess) return BI_HIP_EDEVICE;
        if (e.add_scratch) (void)hipFree(e.add_scratch);
        e.add_scratch = nullptr;
        e.add_bytes = 0;
        if (bi::call_malloc((void**)&e.add_scratch, need) != hipSuccess) return BI_HIP_ENOMEM;
        e.add_bytes = need;
    }
    uint32_t* words = (uint32_t*)(e.add_scratch + desc_bytes);
    uint32_t* flags = words + 2 * tiles;
    size_t off = 0;
    for (size_t i = 0; i < count; ++i) {
        hj[i].maps = words + off;
        hj[i].carry_in = words + tiles + off;
        hj[i].flags = flags + 2 * i;
        off += hj[i].ntiles;
    }
    auto* d_jobs = (bi::AddJob*)e.add_scratch;
    if (hipMemcpyAsync(d_jobs, hj.data(), count * sizeof(bi::AddJob), hipMemcpyHostToDevice, st) != hipSuccess)
        return BI_HIP_EDEVICE;
    if (bi::launch_add_batch(d_jobs, (uint32_t)count, max_tiles, st) != hipSuccess) return BI_HIP_EDEVICE;
    std::vector<uint32_t> hf(2 * count);
    if (hipMemcpyAsync(hf.data(), flags, hf.size() * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return BI_HIP_EDEVICE;
    for (size_t i = 0; i < count; ++i) {
        jobs[i].swapped = (int)hf[2 * i];
        jobs[i].out_len = hf[2 * i + 1];
    }
    return BI_HIP_OK;
}

int bi_hip_device_alloc(size_t bytes, void** ptr) {
    if (!ptr) return BI_HIP_EINVAL;
    *ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(engine().mtx);
        if (int rc = engine().init(0)) return rc;
    }
    if (bytes == 0) return BI_HIP_OK;
    return bi::call_malloc(ptr, bytes) == hipSuccess ? BI_HIP_OK : BI_HIP_ENOMEM;
}

int bi_hip_device_free(void* ptr) {
    if (!ptr) return BI_HIP_OK;
    return hipFree(ptr) == hipSuccess ? BI_HIP_OK : BI_HIP_EDEVICE;
}

int bi_hip_copy_to_device(void* d_dst, const void* const* h_src, const size_t* bytes, size_t count) {
    if ((!h_src || !bytes) && count) return BI_HIP_EINVAL;
    bi::HostSegs h;
    for (size_t i = 0; i < count; ++i) h.add(h_src[i], bytes[i]);
    if (h.total() == 0) return BI_HIP_OK;
    if (!d_dst) return BI_HIP_EINVAL;
    bi::Engine& e = engine();
    {
        std::lock_guard<std::mutex> lk(e.mtx);
        if (int rc = e.init(0)) return rc;
    }
    std::lock_guard<std::mutex> plk(e.pipe_mtx);
    if (e.pipe.h2d(d_dst, h, e.stream) != hipSuccess || hipStreamSynchronize(e.stream) != hipSuccess)
        return BI_HIP_EDEVICE;
    return BI_HIP_OK;
}

int bi_hip_copy_to_host(void* const* h_dst, const size_t* bytes, size_t count, const void* d_src) {
    if ((!h_dst || !bytes) && count) return BI_HIP_EINVAL;
    bi::HostSegs h;
    for (size_t i = 0; i < count; ++i) h.add(h_dst[i], bytes[i]);
    if (h.total() == 0) return BI_HIP_OK;
    if (!d_src) return BI_HIP_EINVAL;
    bi::Engine& e = engine();
    {
        std::lock_guard<std::mutex> lk(e.mtx);
        if (int rc = e.init(0)) return rc;
    }
    std::lock_guard<std::mutex> plk(e.pipe_mtx);
    return e.pipe.d2h(h, d_src, e.stream) == hipSuccess ? BI_HIP_OK : BI_HIP_EDEVICE;
}

int bi_hip_synchronize(void) {
    std::lock_guard<std::mutex> lk(engine().mtx);
    if (!engine().ready) return BI_HIP_OK;
    if (engine().wait_shared()) return BI_HIP_EDEVICE;  // products queued on callers' streams
    return hipStreamSynchronize(engine().stream) == hipSuccess ? BI_HIP_OK : BI_HIP_EDEVICE;
}

const char* bi_hip_strerror(int code) {
    switch (code) {
        case BI_HIP_OK: return "ok";
        case BI_HIP_EINVAL: return "invalid argument";
        case BI_HIP_ETOOBIG: return "product of 2^32 - 8192 limbs or more";
        case BI_HIP_ENOMEM: return "out of memory";
        case BI_HIP_EDEVICE: return "HIP runtime error";
        case BI_HIP_ENODEV: return "no gfx950 device";
        case BI_HIP_ECAP: return "output capacity too small";
        case BI_HIP_ESCRATCH: return "product too large for the allocation-free scratch (transforms up to 2^20 points)";
        case BI_HIP_ENOINIT: return "engine not initialised: call bi_hip_init first";
        default: return "unknown status";
    }
}

}  // extern "C"
