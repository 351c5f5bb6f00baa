// Host side of the NTT multiply engine: root tables, per-length plans, workspace and the C ABI
// declared in include/bigint_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/bigint_hip.h"
#include "ntt_plan.h"

namespace bi {
namespace {

constexpr uint32_t kPh[3] = {469762049u, 2013265921u, 2281701377u};
constexpr uint32_t kGh[3] = {3u, 31u, 3u};
constexpr int kTileH = 4096;
constexpr uint32_t kMaxLog = 26;

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
    uint32_t* beta = nullptr;      // device [3][nblk] (null when nblk == 1)
    uint32_t* beta_inv = nullptr;
};

struct Plan {
    uint32_t log_n = 0;
    std::vector<OuterPass> outer;  // forward order
    uint32_t mid_log_r = 0, mid_nblk = 1;
    uint32_t* mid_beta = nullptr;
    uint32_t* mid_beta_inv = nullptr;
    uint32_t kscale[3] = {0, 0, 0};
};

hipError_t upload_beta(uint32_t log_n, uint32_t log_r, uint32_t log_s, uint32_t nblk,
                       uint32_t** fwd, uint32_t** inv) {
    *fwd = *inv = nullptr;
    if (nblk <= 1) return hipSuccess;
    std::vector<uint32_t> hf(3 * (size_t)nblk), hi(3 * (size_t)nblk);
    const uint32_t lb = 31u - (uint32_t)__builtin_clz(nblk);
    (void)log_r;
    for (int I = 0; I < 3; ++I) {
        const uint32_t p = kPh[I];
        // beta_blk = w_{n/S}^{bitrev_{log nblk}(blk)}
        const uint32_t wf = root(I, log_n - log_s, false), wi = root(I, log_n - log_s, true);
        for (uint32_t b = 0; b < nblk; ++b) {
            const uint32_t e = bitrev(b, lb);
            hf[I * (size_t)nblk + b] = to_mont(powmod(wf, e, p), p);
            hi[I * (size_t)nblk + b] = to_mont(powmod(wi, e, p), p);
        }
    }
    hipError_t e = hipMalloc((void**)fwd, hf.size() * 4);
    if (e != hipSuccess) return e;
    e = hipMalloc((void**)inv, hi.size() * 4);
    if (e != hipSuccess) return e;
    e = hipMemcpy(*fwd, hf.data(), hf.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    return hipMemcpy(*inv, hi.data(), hi.size() * 4, hipMemcpyHostToDevice);
}

struct ProfRec {
    int kind;
    hipEvent_t e0, e1;
    double bytes;
};

class Engine {
public:
    std::mutex mtx;
    bool profiling = false;
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
    // Bracket one launch with events when profiling.
    template <class F>
    hipError_t timed(int kind, double bytes, hipStream_t st, F&& launch) {
        if (!profiling) return launch();
        ProfRec r{kind, get_event(), get_event(), bytes};
        if (!r.e0 || !r.e1) return hipErrorOutOfMemory;
        hipEventRecord(r.e0, st);
        hipError_t e = launch();
        hipEventRecord(r.e1, st);
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
    uint32_t* wtab = nullptr;  // [2][3][4096]
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
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess)
            return BI_HIP_EDEVICE;
        // root table W[m + i] = w_{2m}^{bitrev_{log m}(i)}, Montgomery form
        std::vector<uint32_t> h(2 * 3 * (size_t)kTileH);
        for (int dir = 0; dir < 2; ++dir)
            for (int I = 0; I < 3; ++I) {
                uint32_t* t = h.data() + ((size_t)dir * 3 + I) * kTileH;
                t[0] = to_mont(1, kPh[I]);
                for (uint32_t lm = 0; (1u << lm) < (uint32_t)kTileH; ++lm) {
                    const uint32_t m = 1u << lm;
                    const uint32_t w = root(I, lm + 1, dir == 1);
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
        ready = true;
        return BI_HIP_OK;
    }

    Plan* plan(uint32_t log_n) {
        auto it = plans.find(log_n);
        if (it != plans.end()) return it->second.get();
        auto pl = std::make_unique<Plan>();
        pl->log_n = log_n;
        const uint32_t n = 1u << log_n;
        if (log_n <= 12) {
            pl->mid_log_r = log_n;
            pl->mid_nblk = 1;
        } else {
            const uint32_t q = log_n - 12;
            std::vector<uint32_t> rs;
            if (q <= 8) rs = {q};
            else rs = {(q + 1) / 2, q / 2};
            uint32_t done = 0;
            for (uint32_t r : rs) {
                OuterPass op;
                op.log_r = r;
                op.log_c = 12 - r;
                op.log_s = log_n - done - r;
                op.nblk = 1u << done;
                if (upload_beta(log_n, r, op.log_s, op.nblk, &op.beta, &op.beta_inv) != hipSuccess)
                    return nullptr;
                pl->outer.push_back(op);
                done += r;
            }
            pl->mid_log_r = 12;
            pl->mid_nblk = n >> 12;
            if (upload_beta(log_n, 12, 0, pl->mid_nblk, &pl->mid_beta, &pl->mid_beta_inv) !=
                hipSuccess)
                return nullptr;
        }
        for (int I = 0; I < 3; ++I) {
            const uint32_t p = kPh[I];
            const uint32_t r2 = mulmod(to_mont(1, p), to_mont(1, p), p);  // R^2 mod p
            pl->kscale[I] = mulmod(invmod(n % p, p), r2, p);
        }
        Plan* raw = pl.get();
        plans[log_n] = std::move(pl);
        return raw;
    }

    // device workspace of at least `words` u32
    uint32_t* workspace(size_t words) {
        if (words > ws_words) {
            if (ws) hipFree(ws);
            ws = nullptr;
            ws_words = 0;
            if (hipMalloc((void**)&ws, words * 4) != hipSuccess) return nullptr;
            ws_words = words;
        }
        return ws;
    }

    // Launch the whole pass sequence of one plan for njobs products of the same transform
    // length.  single/gsingle are used when njobs == 1 and no descriptor arrays are given.
    int launch_products(const Plan* pl, bool square, uint32_t njobs, const PassJob& single,
                        const PassJob* d_jobs, uint32_t ngjobs, const GarnerJob& gsingle,
                        const GarnerJob* d_gjobs, uint32_t max_out_len, double in_limbs,
                        double out_limbs, uint32_t* status, uint32_t status_words, hipStream_t st) {
        const uint32_t n = 1u << pl->log_n;
        const double nj = (double)njobs;
        if (timed(BI_HIP_K_FILL, 4.0 * status_words, st,
                  [&] { return launch_fill(status, 0u, status_words, st); }) != hipSuccess)
            return BI_HIP_EDEVICE;
        const double nops_d = square ? 1.0 : 2.0;
        PassArgs a{};
        a.job = single;
        a.jobs = d_jobs;
        a.log_n = pl->log_n;
        a.wtab = wtab;
        const uint32_t sq = square ? (uint32_t)kSquare : 0u;
        const uint32_t nops = square ? 1u : 2u;
        for (size_t k = 0; k < pl->outer.size(); ++k) {
            const OuterPass& op = pl->outer[k];
            a.log_r = op.log_r;
            a.log_c = op.log_c;
            a.log_s = op.log_s;
            a.flags = sq | (k == 0 ? (uint32_t)kFromLimbs : 0u);
            a.beta = op.beta;
            a.beta_inv = nullptr;
            a.beta_stride = op.nblk;
            // the first pass reads the limbs once; every pass writes 3 planes per operand
            const double bytes = k == 0 ? 4.0 * in_limbs + 12.0 * n * nops_d * nj : 24.0 * n * nops_d * nj;
            if (timed(BI_HIP_K_OUTER_FWD, bytes, st,
                      [&] { return launch_outer_pass(a, false, njobs, nops, st); }) != hipSuccess)
                return BI_HIP_EDEVICE;
        }
        a.log_r = pl->mid_log_r;
        a.log_c = 0;
        a.log_s = 0;
        a.flags = sq | (pl->outer.empty() ? (uint32_t)kFromLimbs : 0u);
        a.beta = pl->mid_beta;
        a.beta_inv = pl->mid_beta_inv;
        a.beta_stride = pl->mid_nblk;
        {
            const double in_bytes = pl->outer.empty() ? 4.0 * in_limbs : 12.0 * n * nops_d * nj;
            if (timed(BI_HIP_K_MID, in_bytes + 12.0 * n * nj, st,
                      [&] { return launch_mid_pass(a, njobs, st); }) != hipSuccess)
                return BI_HIP_EDEVICE;
        }
        for (size_t k = pl->outer.size(); k-- > 0;) {
            const OuterPass& op = pl->outer[k];
            a.log_r = op.log_r;
            a.log_c = op.log_c;
            a.log_s = op.log_s;
            a.flags = 0;
            a.beta = op.beta_inv;
            a.beta_inv = nullptr;
            a.beta_stride = op.nblk;
            if (timed(BI_HIP_K_OUTER_INV, 24.0 * n * nj, st,
                      [&] { return launch_outer_pass(a, true, njobs, 1, st); }) != hipSuccess)
                return BI_HIP_EDEVICE;
        }
        GarnerArgs g{};
        g.job = gsingle;
        g.jobs = d_gjobs;
        g.k = gk;
        if (timed(BI_HIP_K_GARNER, 12.0 * n * nj + 4.0 * out_limbs, st,
                  [&] { return launch_garner_carry(g, max_out_len, ngjobs, st); }) != hipSuccess)
            return BI_HIP_EDEVICE;
        return BI_HIP_OK;
    }

    int run(const uint32_t* d_a, uint32_t la, const uint32_t* d_b, uint32_t lb, bool square,
            uint32_t* d_out, hipStream_t st) {
        const int lg = bi_hip_transform_log2(la, lb);
        if (lg < 0) return BI_HIP_ETOOBIG;
        Plan* pl = plan((uint32_t)lg);
        if (!pl) return BI_HIP_ENOMEM;
        const uint32_t n = 1u << lg;
        const uint32_t out_len = la + lb;
        const uint32_t ntiles = (out_len + 4095u) / 4096u;
        const size_t need = (size_t)3 * n * (square ? 1 : 2) + ntiles + 64;
        uint32_t* w = workspace(need);
        if (!w) return BI_HIP_ENOMEM;
        uint32_t* res_a = w;
        uint32_t* res_b = square ? nullptr : w + (size_t)3 * n;
        uint32_t* status = w + (size_t)3 * n * (square ? 1 : 2);
        const PassJob job{d_a, square ? d_a : d_b, la, square ? la : lb, res_a, res_b};
        GarnerJob gj{};
        gj.y = res_a;
        gj.out = d_out;
        gj.n = n;
        gj.out_len = out_len;
        for (int I = 0; I < 3; ++I) gj.kscale[I] = pl->kscale[I];
        gj.status = status;
        gj.counter = status + ntiles;
        return launch_products(pl, square, 1, job, nullptr, 1, gj, nullptr, out_len,
                               (double)la + (square ? 0.0 : (double)lb), (double)out_len, status,
                               ntiles + 1, st);
    }

    // descriptor buffer for batched launches (device), grow-only
    uint8_t* desc = nullptr;
    size_t desc_bytes = 0;

    // Plan of one product: transform length and optional chunking of the longer operand.
    struct PairPlan {
        uint32_t lg;         // log2 of the transform length
        uint32_t log_chunk;  // 0: ordinary product; else chunk length 2^log_chunk, lg = log_chunk+1
    };
    // Default policy: cut the longer operand into 2^c-limb chunks (transform 2^(c+1)) when the
    // operands are unbalanced enough that the chunk transforms are at least 4x shorter.
    PairPlan default_plan(uint32_t la, uint32_t lb, bool sq) const {
        const int lg = bi_hip_transform_log2(la, lb);
        PairPlan p{(uint32_t)lg, 0};
        if (sq || lg < 14) return p;
        const uint32_t lo = std::min(la, lb);
        uint32_t c = 1;
        while ((1u << c) < lo) ++c;  // 2^c >= min(la, lb) >= lb - 1
        if (c < 11) c = 11;
        if (c + 1 + 2 <= (uint32_t)lg) p = PairPlan{c + 1, c};
        return p;
    }

    int plan_mode = BI_HIP_PLAN_TABLE;
    // external per-pair plan override (set by the planner), consulted by run_batch
    std::vector<PairPlan> plan_override;

    // Batched products on device pointers, grouped by transform length and squaring.
    int run_batch(const bi_hip_dpair* pairs, size_t count, hipStream_t st) {
        struct Item {
            uint32_t idx, la, lb;
            bool sq, swap;
            PairPlan pp;
            uint32_t nchunks;
        };
        std::map<std::pair<uint32_t, bool>, std::vector<Item>> groups;
        for (size_t i = 0; i < count; ++i) {
            const bi_hip_dpair& p = pairs[i];
            const bool sq = p.b == nullptr || (p.b == p.a && p.lb == p.la);
            const size_t lb = sq ? p.la : p.lb;
            if (!p.a || p.la == 0 || lb == 0 || !p.out || p.la > 0xFFFFFFFFull || lb > 0xFFFFFFFFull)
                return BI_HIP_EINVAL;
            if (p.out_cap < p.la + lb) return BI_HIP_ECAP;
            const int lg = bi_hip_transform_log2(p.la, lb);
            if (lg < 0) return BI_HIP_ETOOBIG;
            Item it{(uint32_t)i, (uint32_t)p.la, (uint32_t)lb, sq, false, PairPlan{(uint32_t)lg, 0}, 1};
            PairPlan pp = i < plan_override.size() ? plan_override[i] : default_plan(it.la, it.lb, sq);
            if (pp.log_chunk && !sq) {
                it.swap = it.la < it.lb;  // chunk the longer operand
                const uint32_t la_ = std::max(it.la, it.lb), lb_ = std::min(it.la, it.lb);
                const uint32_t L = 1u << pp.log_chunk;
                if (pp.lg == pp.log_chunk + 1 && lb_ <= L + 1 && pp.lg <= 26) {
                    it.pp = pp;
                    it.nchunks = (la_ + L - 1) / L;
                }
            } else if (pp.lg >= (uint32_t)lg && pp.lg <= 26) {
                it.pp = PairPlan{pp.lg, 0};
            }
            groups[{it.pp.lg, sq}].push_back(it);
        }
        size_t ws_need = 0, desc_need = 0;
        for (auto& [key, items] : groups) {
            const size_t n = (size_t)1 << key.first;
            size_t words = 0, njobs = 0;
            for (auto& it : items) {
                words += 3 * n * (it.sq ? 1 : 2) * it.nchunks + (it.la + it.lb + 4095u) / 4096u + 1;
                njobs += it.nchunks;
            }
            ws_need = std::max(ws_need, words + 64);
            desc_need += njobs * sizeof(PassJob) + items.size() * sizeof(GarnerJob) + 512;
        }
        uint32_t* w = workspace(ws_need);
        if (!w) return BI_HIP_ENOMEM;
        if (desc_need > desc_bytes) {
            if (desc) hipFree(desc);
            desc = nullptr;
            desc_bytes = 0;
            if (hipMalloc((void**)&desc, desc_need) != hipSuccess) return BI_HIP_ENOMEM;
            desc_bytes = desc_need;
        }
        std::vector<uint8_t> hdesc(desc_need);
        struct Launch {
            Plan* pl;
            bool sq;
            uint32_t njobs, ngjobs, max_out;
            size_t pj_off, gj_off;
            double in_limbs, out_limbs;
            uint32_t status_words;
            uint32_t* status;
        };
        std::vector<Launch> launches;
        size_t off = 0;
        for (auto& [key, items] : groups) {
            Plan* pl = plan(key.first);
            if (!pl) return BI_HIP_ENOMEM;
            const size_t n = (size_t)1 << key.first;
            Launch L{pl, key.second, 0, (uint32_t)items.size(), 0, 0, 0, 0.0, 0.0, 0, nullptr};
            uint32_t njobs = 0;
            for (auto& it : items) njobs += it.nchunks;
            L.njobs = njobs;
            L.pj_off = off;
            off += njobs * sizeof(PassJob);
            off = (off + 255) & ~(size_t)255;
            L.gj_off = off;
            off += items.size() * sizeof(GarnerJob);
            off = (off + 255) & ~(size_t)255;
            auto* pj = reinterpret_cast<PassJob*>(hdesc.data() + L.pj_off);
            auto* gj = reinterpret_cast<GarnerJob*>(hdesc.data() + L.gj_off);
            size_t res_words = 0;
            for (auto& it : items) res_words += 3 * n * (it.sq ? 1 : 2) * it.nchunks;
            uint32_t* res = w;
            uint32_t* status = w + res_words;
            uint32_t* st_cur = status;
            size_t jp = 0;
            for (size_t j = 0; j < items.size(); ++j) {
                const Item& it = items[j];
                const bi_hip_dpair& p = pairs[it.idx];
                const uint32_t* A = it.swap ? p.b : p.a;
                const uint32_t* B = it.sq ? p.a : (it.swap ? p.a : p.b);
                const uint32_t la = it.swap ? it.lb : it.la, lb = it.swap ? it.la : it.lb;
                GarnerJob g{};
                g.y = res;
                if (it.nchunks > 1 || it.pp.log_chunk) {
                    const uint32_t Lc = 1u << it.pp.log_chunk;
                    for (uint32_t c = 0; c < it.nchunks; ++c) {
                        const uint32_t lac = std::min(Lc, la - c * Lc);
                        pj[jp++] = PassJob{A + (size_t)c * Lc, B, lac, lb, res, res + 3 * n};
                        res += 6 * n;
                    }
                    g.nchunks = it.nchunks;
                    g.log_chunk = it.pp.log_chunk;
                    g.chunk_stride = 6 * n;
                } else {
                    uint32_t* ra = res;
                    res += 3 * n;
                    uint32_t* rb = nullptr;
                    if (!it.sq) {
                        rb = res;
                        res += 3 * n;
                    }
                    pj[jp++] = PassJob{A, B, la, lb, ra, rb};
                    g.nchunks = 1;
                }
                const uint32_t out_len = it.la + it.lb;
                const uint32_t nt = (out_len + 4095u) / 4096u;
                g.out = p.out;
                g.n = (uint32_t)n;
                g.out_len = out_len;
                for (int I = 0; I < 3; ++I) g.kscale[I] = pl->kscale[I];
                g.status = st_cur;
                g.counter = st_cur + nt;
                st_cur += nt + 1;
                gj[j] = g;
                L.max_out = std::max(L.max_out, out_len);
                L.in_limbs += (double)it.la + (it.sq ? 0.0 : (double)it.lb);
                L.out_limbs += (double)out_len;
            }
            L.status_words = (uint32_t)(st_cur - status);
            L.status = status;
            launches.push_back(L);
        }
        if (hipStreamSynchronize(st) != hipSuccess) return BI_HIP_EDEVICE;
        if (hipMemcpy(desc, hdesc.data(), off, hipMemcpyHostToDevice) != hipSuccess)
            return BI_HIP_EDEVICE;
        for (auto& L : launches) {
            const auto* pj = reinterpret_cast<const PassJob*>(desc + L.pj_off);
            const auto* gj = reinterpret_cast<const GarnerJob*>(desc + L.gj_off);
            const auto* hp = reinterpret_cast<const PassJob*>(hdesc.data() + L.pj_off);
            const auto* hg = reinterpret_cast<const GarnerJob*>(hdesc.data() + L.gj_off);
            const int rc = launch_products(L.pl, L.sq, L.njobs, hp[0], L.njobs == 1 ? nullptr : pj,
                                           L.ngjobs, hg[0], L.ngjobs == 1 ? nullptr : gj, L.max_out,
                                           L.in_limbs, L.out_limbs, L.status, L.status_words, st);
            if (rc) return rc;
        }
        return BI_HIP_OK;
    }
};

Engine& engine() {
    static Engine e;
    return e;
}

int check_operand(const uint32_t* a, size_t la) {
    if (!a || la == 0) return BI_HIP_EINVAL;
    if (la > 0xFFFFFFFFull) return BI_HIP_ETOOBIG;
    return BI_HIP_OK;
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
    if (la == 0 || lb == 0) return -1;
    const size_t need = la + lb - 1;
    if (need > ((size_t)1 << bi::kMaxLog)) return -1;
    int lg = 4;  // at least one 16-point tile
    while (((size_t)1 << lg) < need) ++lg;
    return lg;
}

int bi_hip_mul_device(const uint32_t* d_a, size_t la, const uint32_t* d_b, size_t lb,
                      int is_square, uint32_t* d_out, size_t out_cap, void* stream) {
    if (bi::check_operand(d_a, la) || (!is_square && bi::check_operand(d_b, lb)) || !d_out)
        return BI_HIP_EINVAL;
    if (is_square) lb = la;
    if (out_cap < la + lb) return BI_HIP_ECAP;
    std::lock_guard<std::mutex> lk(engine().mtx);
    int rc = engine().init(0);
    if (rc) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : engine().stream;
    return engine().run(d_a, (uint32_t)la, is_square ? d_a : d_b, (uint32_t)lb, is_square != 0,
                        d_out, st);
}

int bi_hip_mul(const uint32_t* a, size_t la, const uint32_t* b, size_t lb, int is_square,
               uint32_t* out, size_t out_cap, size_t* out_len) {
    if (bi::check_operand(a, la) || (!is_square && bi::check_operand(b, lb)) || !out || !out_len)
        return BI_HIP_EINVAL;
    if (is_square) {
        b = a;
        lb = la;
    }
    if (a[la - 1] == 0 || b[lb - 1] == 0) return BI_HIP_EINVAL;
    if (out_cap < la + lb) return BI_HIP_ECAP;
    if (bi_hip_transform_log2(la, lb) < 0) return BI_HIP_ETOOBIG;
    std::lock_guard<std::mutex> lk(engine().mtx);
    bi::Engine& e = engine();
    int rc = e.init(0);
    if (rc) return rc;
    const size_t words = la + (is_square ? 0 : lb) + la + lb;
    if (words > e.io_words) {
        if (e.io) hipFree(e.io);
        e.io = nullptr;
        e.io_words = 0;
        if (hipMalloc((void**)&e.io, words * 4) != hipSuccess) return BI_HIP_ENOMEM;
        e.io_words = words;
    }
    uint32_t* da = e.io;
    uint32_t* db = is_square ? da : da + la;
    uint32_t* dout = da + la + (is_square ? 0 : lb);
    if (hipMemcpyAsync(da, a, la * 4, hipMemcpyHostToDevice, e.stream) != hipSuccess)
        return BI_HIP_EDEVICE;
    if (!is_square &&
        hipMemcpyAsync(db, b, lb * 4, hipMemcpyHostToDevice, e.stream) != hipSuccess)
        return BI_HIP_EDEVICE;
    rc = e.run(da, (uint32_t)la, db, (uint32_t)lb, is_square != 0, dout, e.stream);
    if (rc) return rc;
    if (hipMemcpyAsync(out, dout, (la + lb) * 4, hipMemcpyDeviceToHost, e.stream) != hipSuccess)
        return BI_HIP_EDEVICE;
    if (hipStreamSynchronize(e.stream) != hipSuccess) return BI_HIP_EDEVICE;
    size_t n = la + lb;
    while (n > 1 && out[n - 1] == 0) --n;
    *out_len = n;
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
            rc = e.run(p.a, (uint32_t)p.la, sq ? p.a : p.b, (uint32_t)lb, sq, p.out, st);
            if (rc) return rc;
        }
        return BI_HIP_OK;
    }
    return e.run_batch(pairs, count, st);
}

int bi_hip_mul_batch(bi_hip_pair* pairs, size_t count, int plan_mode) {
    if (!pairs && count) return BI_HIP_EINVAL;
    if (count == 0) return BI_HIP_OK;
    // validate, stage every operand in one device buffer, run the device batch, copy back
    size_t in_words = 0, out_words = 0;
    for (size_t i = 0; i < count; ++i) {
        bi_hip_pair& p = pairs[i];
        const bool sq = p.a == p.b && p.la == p.lb;
        if (bi::check_operand(p.a, p.la) || (!sq && bi::check_operand(p.b, p.lb)) || !p.out) {
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
        if (bi_hip_transform_log2(p.la, p.lb) < 0) {
            p.status = BI_HIP_ETOOBIG;
            return BI_HIP_ETOOBIG;
        }
        in_words += p.la + (sq ? 0 : p.lb);
        out_words += p.la + p.lb;
    }
    uint32_t* dbuf = nullptr;
    {
        std::lock_guard<std::mutex> lk(engine().mtx);
        int rc = engine().init(0);
        if (rc) return rc;
    }
    if (hipMalloc((void**)&dbuf, (in_words + out_words) * 4) != hipSuccess) return BI_HIP_ENOMEM;
    std::vector<uint32_t> hin(in_words);
    std::vector<bi_hip_dpair> dp(count);
    size_t io = 0, oo = in_words;
    for (size_t i = 0; i < count; ++i) {
        bi_hip_pair& p = pairs[i];
        const bool sq = p.a == p.b && p.la == p.lb;
        std::memcpy(hin.data() + io, p.a, p.la * 4);
        dp[i].a = dbuf + io;
        dp[i].la = p.la;
        io += p.la;
        if (sq) {
            dp[i].b = dp[i].a;
        } else {
            std::memcpy(hin.data() + io, p.b, p.lb * 4);
            dp[i].b = dbuf + io;
            io += p.lb;
        }
        dp[i].lb = p.lb;
        dp[i].out = dbuf + oo;
        dp[i].out_cap = p.la + p.lb;
        oo += p.la + p.lb;
    }
    int rc = BI_HIP_OK;
    if (hipMemcpy(dbuf, hin.data(), in_words * 4, hipMemcpyHostToDevice) != hipSuccess) rc = BI_HIP_EDEVICE;
    if (!rc) rc = bi_hip_mul_batch_device(dp.data(), count, plan_mode, nullptr);
    if (!rc) rc = bi_hip_synchronize();
    std::vector<uint32_t> hout;
    if (!rc) {
        hout.resize(out_words);
        if (hipMemcpy(hout.data(), dbuf + in_words, out_words * 4, hipMemcpyDeviceToHost) != hipSuccess)
            rc = BI_HIP_EDEVICE;
    }
    hipFree(dbuf);
    size_t o = 0;
    for (size_t i = 0; i < count; ++i) {
        bi_hip_pair& p = pairs[i];
        p.status = rc;
        if (rc) continue;
        const size_t len = p.la + p.lb;
        std::memcpy(p.out, hout.data() + o, len * 4);
        o += len;
        size_t n = len;
        while (n > 1 && p.out[n - 1] == 0) --n;
        p.out_len = n;
    }
    return rc;
}

int bi_hip_profile(int enable) {
    std::lock_guard<std::mutex> lk(engine().mtx);
    bi::Engine& e = engine();
    if (enable) {
        if (e.ready) hipStreamSynchronize(e.stream);
        e.prof_clear();
    }
    e.profiling = enable != 0;
    return BI_HIP_OK;
}

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

int bi_hip_synchronize(void) {
    std::lock_guard<std::mutex> lk(engine().mtx);
    if (!engine().ready) return BI_HIP_OK;
    return hipStreamSynchronize(engine().stream) == hipSuccess ? BI_HIP_OK : BI_HIP_EDEVICE;
}

const char* bi_hip_strerror(int code) {
    switch (code) {
        case BI_HIP_OK: return "ok";
        case BI_HIP_EINVAL: return "invalid argument";
        case BI_HIP_ETOOBIG: return "product needs a transform longer than 2^26 points";
        case BI_HIP_ENOMEM: return "out of memory";
        case BI_HIP_EDEVICE: return "HIP runtime error";
        case BI_HIP_ENODEV: return "no gfx950 device";
        case BI_HIP_ECAP: return "output capacity too small";
        default: return "unknown status";
    }
}

}  // extern "C"
