/* Test infrastructure for large batched workloads: the seeded operand generator and residue
 * checks of bigint_amd/limbs.py in C, so tests can build and check gigabit batches in seconds.
 *
 * chk_random_operand(bits, seed, out): the operand limbs.random_operand(bits, seed) returns:
 *   ceil(bits/64) SplitMix64 words from `seed`, masked to `bits` bits with bit bits-1 set, cut
 *   into ceil(bits/28) 28-bit limbs.
 * chk_residues(limbs, off, len, count, q, out): out[i] = value of limbs[off[i] .. off[i]+len[i])
 *   (28-bit little-endian limbs) mod q, q < 2^31; pairs are spread over threads.
 * Used only by tests/ as the checker.  Nothing under bigint_amd/ links or calls it.
 */
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

static uint64_t splitmix_next(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

size_t chk_random_operand(uint64_t bits, uint64_t seed, uint32_t* out) {
    if (bits == 0) return 0;
    const uint64_t nwords = (bits + 63) / 64, nlimbs = (bits + 27) / 28;
    uint64_t state = seed;
    unsigned __int128 acc = 0;
    int have = 0;
    uint64_t li = 0;
    for (uint64_t i = 0; i < nwords; ++i) {
        uint64_t x = splitmix_next(&state);
        if (i == nwords - 1) {
            const unsigned r = (unsigned)(bits % 64);
            if (r) x &= (1ull << r) - 1;
            x |= 1ull << ((bits - 1) % 64);
        }
        acc |= (unsigned __int128)x << have;
        have += 64;
        while (have >= 28 && li < nlimbs) {
            out[li++] = (uint32_t)acc & 0xFFFFFFFu;
            acc >>= 28;
            have -= 28;
        }
    }
    while (li < nlimbs) {
        out[li++] = (uint32_t)acc & 0xFFFFFFFu;
        acc >>= 28;
    }
    return (size_t)nlimbs;
}

/* Barrett reduction of x < 2^60 by q < 2^31 with m = floor(2^64 / q). */
static inline uint64_t barrett(uint64_t x, uint64_t q, uint64_t m) {
    const uint64_t t = (uint64_t)(((unsigned __int128)x * m) >> 64);
    uint64_t r = x - t * q;
    if (r >= q) r -= q;
    if (r >= q) r -= q;
    return r;
}

struct res_job {
    const uint32_t* limbs;
    const uint64_t* off;
    const uint64_t* len;
    size_t lo, hi;
    uint64_t q;
    uint32_t* out;
};

static void* res_worker(void* p) {
    const struct res_job* j = (const struct res_job*)p;
    const uint64_t q = j->q, m = (uint64_t)(((unsigned __int128)1 << 64) / q);
    for (size_t i = j->lo; i < j->hi; ++i) {
        const uint32_t* x = j->limbs + j->off[i];
        uint64_t r = 0;
        for (uint64_t k = j->len[i]; k-- > 0;) r = barrett((r << 28) | x[k], q, m);  /* Horner, top down */
        j->out[i] = (uint32_t)r;
    }
    return NULL;
}

int chk_residues(const uint32_t* limbs, const uint64_t* off, const uint64_t* len, size_t count,
                 uint32_t q, uint32_t* out, int nthreads) {
    if (q < 2 || q >= 0x80000000u) return 1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    /* pairs split so each thread gets about the same number of limbs */
    uint64_t total = 0;
    for (size_t i = 0; i < count; ++i) total += len[i];
    pthread_t th[64];
    struct res_job jobs[64];
    size_t start = 0;
    int started = 0;
    uint64_t acc = 0;
    for (int t = 0; t < nthreads && start < count; ++t) {
        const uint64_t goal = total * (uint64_t)(t + 1) / (uint64_t)nthreads;
        size_t end = start;
        while (end < count && (acc < goal || end == start)) acc += len[end++];
        if (t == nthreads - 1) {
            while (end < count) acc += len[end++];
        }
        jobs[t] = (struct res_job){limbs, off, len, start, end, q, out};
        if (pthread_create(&th[t], NULL, res_worker, &jobs[t]) != 0) {
            res_worker(&jobs[t]);
            th[t] = 0;
        }
        ++started;
        start = end;
    }
    for (int t = 0; t < started; ++t)
        if (th[t]) pthread_join(th[t], NULL);
    return 0;
}
