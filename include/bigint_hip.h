/*
 * bigint_hip.h — C ABI of the MI355X (gfx950) NTT multiply engine.
 *
 * Plain C: pointers, sizes and status codes only.  Limbs are the reference's Digits layout:
 * uint32_t, little-endian, each below 2^28 (include/bigint/bigint_base.h:21-26).
 *
 * The engine replaces, one for one, the reference's multiply seam below BigInt/BigFloat:
 *   bi_hip_mul          <- bigint::mul::ntt::mul(span<const u64>, span<const u64>)
 *                          (include/bigint/mul.h:194, src/mul/ntt_multithread.cpp:444-530), taking
 *                          28-bit limbs directly instead of the u64 repack of mul.h:245-300
 *   bi_hip_mul_device   <- the same product with device-resident operands (no reference
 *                          counterpart: the reference has no device memory)
 *   bi_hip_mul_batch    <- a loop of `operator*` calls at one product-tree level
 *                          (examples/pi.cpp:45-58, examples/binary_split.cpp:23-75), done as one
 *                          batched launch sequence
 *   bi_hip_init         <- first-use initialisation done lazily by get_pool()
 *                          (src/mul/ntt_multithread.cpp:287-290)
 * Preconditions follow the reference: operands are nonzero with a nonzero top limb
 * (mul.h:222, 232).  Where the reference hits unreachable() (bigint_base.h:40-46) this ABI
 * returns BI_HIP_EINVAL instead.  Products beyond one 2^26-point transform (ntt_base.h:22-27; the
 * reference sends them to SSA, mul.h:199-219) are computed from chunked transforms, and past a
 * shorter operand of 2^26 words from pieces of it added on the device.  Such products first wait
 * on the host for the work already queued on the stream (the pieces also for their own work, before
 * returning).  Only a product of 2^32 - 8192 limbs or more returns BI_HIP_ETOOBIG.
 *
 * Thread safety: every entry point may be called from any thread; calls are serialised by an
 * internal mutex (the reference serialises pool phases the same way,
 * src/mul/ntt_multithread.cpp:259-260).  The asynchronous entry points share one device
 * workspace; a call queued on another stream than the previous call waits on the device for the
 * previous call's work (hipStreamWaitEvent), so products queued on several streams without a
 * host synchronisation in between are computed one after the other and stay exact.
 *
 * Limbs must be below 2^28 (the Digits invariant).  This is not checked: a larger limb gives a
 * wrong product but never an access outside the operands.
 */
#ifndef BIGINT_HIP_H
#define BIGINT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    BI_HIP_OK = 0,
    BI_HIP_EINVAL = 1,   /* bad argument: null pointer, zero length, top limb zero (host paths) */
    BI_HIP_ETOOBIG = 2,  /* product of 2^32 - 8192 limbs or more */
    BI_HIP_ENOMEM = 3,   /* device or host allocation failed */
    BI_HIP_EDEVICE = 4,  /* HIP runtime error */
    BI_HIP_ENODEV = 5,   /* no gfx950 device */
    BI_HIP_ECAP = 6,     /* output capacity too small */
    BI_HIP_ESCRATCH = 7, /* bi_hip_mul_device_noalloc: product beyond its fixed scratch */
    BI_HIP_ENOINIT = 8   /* bi_hip_mul_device_noalloc: bi_hip_init has not run */
};

/* Plan modes for bi_hip_mul_batch. */
enum {
    BI_HIP_PLAN_FIXED = 0,   /* one launch sequence per product, plan 0 (no chunking) */
    BI_HIP_PLAN_TABLE = 1,   /* autotuned plan table */
    BI_HIP_PLAN_LEARNED = 2  /* learned cost model, table as fallback */
};

/* One product of a batch.  out must hold la + lb limbs; out_len and status are written. */
typedef struct bi_hip_pair {
    const uint32_t* a;
    size_t la;
    const uint32_t* b; /* may equal a (with lb == la) for a square */
    size_t lb;
    uint32_t* out;
    size_t out_cap;
    size_t out_len;
    int status;
} bi_hip_pair;

/* Select the device and build the root tables.  Idempotent; later calls with another device
 * are ignored (first call wins, like init_thread_pool, include/bigint/bigint.h:35-38). */
int bi_hip_init(int device);

/* out = a * b for host limb arrays.  is_square != 0 means b is a (the reference detects this by
 * address, include/bigint/mul.h:225).  *out_len receives the product length without leading zero
 * limbs (la + lb or la + lb - 1).  out_cap must be >= la + lb. */
int bi_hip_mul(const uint32_t* a, size_t la, const uint32_t* b, size_t lb, int is_square,
               uint32_t* out, size_t out_cap, size_t* out_len);

/* The top limbs of a * b for host limb arrays: limbs [la + lb - k, la + lb) of the product go to
 * out[0 .. k) with k = min(keep, la + lb), keep >= 1, and *out_len receives the length of the
 * whole product without leading zero limbs.  Serves precision-truncated products (the
 * reference's mul_digits with output_precision, include/bigint/mul.h:227-243, 302-312), whose
 * low limbs are dropped: only the kept limbs cross to the host. */
int bi_hip_mul_top(const uint32_t* a, size_t la, const uint32_t* b, size_t lb, int is_square,
                   size_t keep, uint32_t* out, size_t* out_len);

/* Device-resident variant: a, b, out are device pointers; exactly la + lb limbs are written
 * (the top one may be zero).  Asynchronous on `stream` (a hipStream_t; NULL = the engine's
 * stream); a chunked plan goes through the batch path (below).  Both single-product calls use the
 * engine's plan choice (the installed table, else the built-in cost heuristic). */
int bi_hip_mul_device(const uint32_t* d_a, size_t la, const uint32_t* d_b, size_t lb,
                      int is_square, uint32_t* d_out, size_t out_cap, void* stream);

/* bi_hip_mul_device without any allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments and result as bi_hip_mul_device, but always the
 * unchunked plan 0, no profiling, and no hipMalloc / hipFree / event creation in the call: the
 * residues and carry status live in a fixed device array of 24 MiB (6 * 2^20 + 1024 words) per
 * device, set up by bi_hip_init (BI_HIP_ENOINIT before that), so a product must fit a transform
 * of at most 2^20 points, about 2^20 words per operand in a balanced product (BI_HIP_ESCRATCH
 * beyond).  Calls share that array: calls that overlap on one GPU (on different streams without
 * ordering between them) race; calls on one stream, or ordered by the caller, do not.  The
 * array keeps the last call's contents; every call clears the words it needs cleared with a
 * kernel (no memset). */
int bi_hip_mul_device_noalloc(const uint32_t* d_a, size_t la, const uint32_t* d_b, size_t lb,
                              int is_square, uint32_t* d_out, size_t out_cap, void* stream);

/* One product of a device-resident batch: a, b, out are device pointers.  b == NULL or
 * (b == a, lb == la) selects squaring.  Exactly la + lb limbs are written to out. */
typedef struct bi_hip_dpair {
    const uint32_t* a;
    size_t la;
    const uint32_t* b;
    size_t lb;
    uint32_t* out;
    size_t out_cap;
} bi_hip_dpair;

/* One signed-magnitude addition of a device-resident batch: x, y, out are device pointers to
 * 28-bit limbs (top limbs may be zero).  subtract == 0: out = |x| + |y| (out_cap >= max(lx, ly)
 * + 1); else out = the larger magnitude minus the smaller (out_cap >= max(lx, ly)), with
 * swapped = 1 when |y| > |x|.  out_len receives the result length without leading zero limbs (0
 * for zero). */
typedef struct bi_hip_dadd {
    const uint32_t* x;
    size_t lx;
    const uint32_t* y;
    size_t ly;
    int subtract;
    uint32_t* out;
    size_t out_cap;
    int swapped;
    size_t out_len;
} bi_hip_dadd;

/* Batched additions on `stream` (NULL = the engine's stream); returns once swapped and out_len
 * are known.  For product-tree levels kept in HBM, whose merge P = P_l Q_r + R_l P_r adds two
 * products (examples/pi.cpp:45-58 does it with BigInt operator+, include/bigint/bigint.h, and
 * src/bigint.cpp:267-548 on the host). */
int bi_hip_add_batch_device(bi_hip_dadd* jobs, size_t count, void* stream);

/* Device memory for callers that keep numbers in HBM between batches (a product tree):
 * hipMalloc / hipFree on the engine's device, and copies through the engine's pinned buffers
 * that gather host ranges into (or scatter device bytes out to) one contiguous device range.
 * Both copies are synchronous.  No counterpart in the reference (its numbers live in host
 * memory). */
int bi_hip_device_alloc(size_t bytes, void** ptr);
int bi_hip_device_free(void* ptr);
int bi_hip_copy_to_device(void* d_dst, const void* const* h_src, const size_t* bytes, size_t count);
int bi_hip_copy_to_host(void* const* h_dst, const size_t* bytes, size_t count, const void* d_src);

/* Device-resident batch, asynchronous on `stream` (NULL = the engine's stream).
 * BI_HIP_PLAN_FIXED runs one launch sequence per product; the other modes group products of
 * equal transform length into one launch sequence per group.  The call does not wait for
 * earlier batches: its descriptors are copied to the device in stream order after the earlier
 * work that reads the descriptor buffer, so batches queued back to back overlap one's host
 * preparation with the previous one's device work.  The host may wait for a batch queued two
 * calls earlier (two staging buffers and two executable graphs are used in turn). */
int bi_hip_mul_batch_device(const bi_hip_dpair* pairs, size_t count, int plan_mode, void* stream);

/* Grouped batches go to the GPU as one captured hipGraph per call (kernels only); two
 * executable graphs are kept and updated in place when a later batch has the same launch shape:
 * enable != 0 (the default) or plain stream launches (0).  No reference counterpart (its product-tree levels are loops of operator*,
 * examples/pi.cpp:45-58). */
int bi_hip_set_graphs(int enable);

/* Batch of independent products (host pointers).  Each pair gets its own status; the call
 * returns the first nonzero status or BI_HIP_OK. */
int bi_hip_mul_batch(bi_hip_pair* pairs, size_t count, int plan_mode);

/* Wait for all work queued by the engine, on its own stream and on callers' streams. */
int bi_hip_synchronize(void);

/* Transform length (log2) of an unchunked la x lb product, or -1 if too large.  The transforms
 * run over the operands' 32-bit words, wa = ceil(7 la / 8) of them (8 limbs = 7 words), and the
 * length is the smallest power of two >= max(wa + wb - 2, wa, wb), at least 16: when
 * wa + wb - 1 exceeds it by one, the top coefficient wraps onto coefficient 0 and the Garner
 * kernel separates the two.  The limit is 2^26 (the primes' 2-adic order, ntt_base.h:22-27):
 * up to 2^25 + 1 words (about 1.07e9 bits) per operand. */
int bi_hip_transform_log2(size_t la, size_t lb);

/* Per-kernel-class timing, for benchmarks: while enabled, every engine launch is bracketed by
 * HIP events on the stream it runs on, and its algorithmic HBM bytes are accumulated. */
enum {
    BI_HIP_K_FILL = 0,      /* status/counter clearing: done by the first kernel of each launch
                               sequence, so no launch of this class is recorded */
    BI_HIP_K_OUTER_FWD = 1, /* strided forward passes after the first */
    BI_HIP_K_MID = 2,       /* contiguous block pass: forward + pointwise + inverse */
    BI_HIP_K_OUTER_INV = 3, /* strided inverse passes */
    BI_HIP_K_GARNER = 4,    /* Garner CRT + carry scan (and the carry fix-up check) */
    BI_HIP_K_OUTER_FIRST = 5, /* first strided forward pass (reads limbs, all three primes) */
    BI_HIP_K_COUNT = 6
};
typedef struct bi_hip_kstat {
    double ms;       /* summed event time */
    double bytes;    /* summed algorithmic HBM bytes */
    uint64_t launches;
} bi_hip_kstat;

/* enable != 0 clears the counters and starts recording; 0 stops. */
int bi_hip_profile(int enable);
/* As bi_hip_profile, but only launches of the kernel classes whose bit (1 << BI_HIP_K_*) is set
 * in mask are bracketed with events, so the other launches run without them; 0 stops. */
int bi_hip_profile_kinds(uint32_t mask);
/* Waits for the recorded launches and copies BI_HIP_K_COUNT entries into stats. */
int bi_hip_profile_read(bi_hip_kstat* stats, size_t cap);

/* ---- plan selection ------------------------------------------------------------------------
 * A plan fixes, for one product: the transform length, chunking of the longer operand, and the
 * block-pass size (hence the radix split of the strided stages).  Plan ids:
 *   0, 1, 2  unchunked, 4096 / 2048 / 1024-point block pass
 *   3, 4, 5  longer operand cut into 2^c, 2^(c+1), 2^(c+2)-limb chunks (2^c = smallest power of two
 *            >= the shorter operand)
 *   6, 7     longer operand cut into 2 / 3 chunks (lengths multiples of the 3584-word carry tile)
 *   8        both operands cut into 2^25-word chunks on 2^26-point transforms, chunk-pair spectra
 *            summed per diagonal (outside the learned plan space; used for products beyond one
 *            transform that no chunk plan fits, and accepted as an explicit id), for a shorter
 *            operand of at most 2^26 words; longer ones run in pieces of that size
 * A chunked product transforms the shorter operand once and multiplies every chunk with that
 * transform; chunk plans are valid only where their transform is shorter than the unchunked one.
 * These replace the reference's hand-set size thresholds (include/bigint/mul.h:80, 153, 192). */
int bi_hip_plan_count(void);
/* 1 if plan_id is valid for an la x lb product (is_square: b is a), else 0. */
int bi_hip_plan_allowed(size_t la, size_t lb, int is_square, int plan_id);
/* Transform-count cost of a plan (transforms x n log2 n; -1 if the plan does not apply): the
 * heuristic's measure, and the baseline the learned model predicts time relative to. */
double bi_hip_plan_cost(size_t la, size_t lb, int is_square, int plan_id);
/* Device batch with an explicit plan id per pair (-1 = the engine's heuristic). */
int bi_hip_mul_batch_device_planned(const bi_hip_dpair* pairs, size_t count,
                                    const int32_t* plan_ids, void* stream);
/* Plan ids the engine would pick under `mode` for shapes[i] = {la, lb, batch}. */
int bi_hip_choose_plans(const uint32_t* shapes, const int* is_square, size_t count, int mode,
                        int32_t* ids);
/* Install the autotuned table (int8 [19][19][13] plan ids indexed by round(log2 bits of a) - 10,
 * round(log2 bits of b) - 10, round(log2 batch), -1 where the cell defers to the built-in cost
 * heuristic; NULL keeps the current one) and/or the MLP (fp32 parameters for K outputs, input
 * normalisation {mean[3], std[3]}; NULL keeps it).  In BI_HIP_PLAN_LEARNED mode the model's
 * choice replaces the table's (or the heuristic's) plan only where it predicts a gain of more
 * than 0.04 in log2 time and the table's plan is predicted to take at least 200 us for the batch
 * (below that, launch costs dominate and the measured table decides). */
int bi_hip_set_planner(const int8_t* table, const float* mlp_params, int K, const float* norm6);
/* Install table and MLP from the binary file bigint_amd/planner.py writes
 * (b"BIPL", int32 K, int8 table[19*19*13], int32 count, float32 params[count], float32 norm[6]);
 * BI_HIP_EINVAL if the file is missing or malformed.  No reference counterpart (the reference's
 * thresholds are compiled in, include/bigint/mul.h:80, 153, 192). */
int bi_hip_load_planner(const char* path);

/* MLP cost model kernels (training and inference on the GPU; device pointers). */
int bi_hip_mlp_param_count(int K);
int bi_hip_mlp_forward(const float* d_x, int B, const float* d_params, int K, float* d_y,
                       void* stream);
/* Masked-MSE loss and gradient; d_grad holds param_count + 1 floats (gradient, then loss). */
int bi_hip_mlp_grad(const float* d_x, const float* d_t, const float* d_mask, int B,
                    const float* d_params, int K, float* d_grad, void* stream);
int bi_hip_mlp_adam(float* d_params, const float* d_grad, float* d_m, float* d_v, int nparams,
                    float lr, float beta1, float beta2, float eps, int step, void* stream);

/* Human-readable text for a status code. */
const char* bi_hip_strerror(int code);

#ifdef __cplusplus
}
#endif

#endif /* BIGINT_HIP_H */
