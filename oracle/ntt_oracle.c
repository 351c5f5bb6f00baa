/*
 * CPU oracle for the three-prime NTT multiply path — TEST INFRASTRUCTURE ONLY.
 *
 * This file is a scalar, single-threaded restatement in plain C of the reference's
 * NTT multiply (Looading99/BigInt).  It exists so the test suite, the smoke check and the
 * CPU-baseline leg of bench.py have an independent checker for the HIP engine.  Nothing in
 * the product (bigint_amd/) may link or call it.
 *
 * What it follows (reference file:line):
 *   - constants P, G, digit width, max length ..... include/bigint/ntt_base.h:14-35
 *   - Montgomery REDC (w = 32), R^2, R mod P ....... include/bigint/ntt_base.h:40-92
 *   - twiddle steps[rev][level] .................... include/bigint/ntt_base.h:195-208
 *   - inv_pow_of_two ............................... include/bigint/ntt_base.h:210-222
 *   - garner_merge ................................. include/bigint/ntt_base.h:228-246
 *   - multi_bit_swap ............................... include/bigint/ntt_base.h:268-287
 *   - imm_mul_num / imm_mul_vec .................... src/mul/ntt_multithread.cpp:292-328
 *   - imm_ntt (radix-2 DIT, cur *= step) ........... src/mul/ntt_multithread.cpp:333-379
 *   - crt_merge .................................... src/mul/ntt_multithread.cpp:384-403
 *   - trim (3-way overlapped sum + carry) .......... src/mul/ntt_multithread.cpp:418-442
 *   - ntt::mul pipeline incl. u64 <-> 28-bit repack  src/mul/ntt_multithread.cpp:444-530
 *   - brute::mul schoolbook ........................ include/bigint/mul.h:78-104
 *   - _mul dispatch thresholds ..................... include/bigint/mul.h:208-220
 *   - mul_digits (precision / zero stripping) ...... include/bigint/mul.h:223-318
 *
 * Deviation: _mul routes the reference's FFT band (16,384 < bits <= 10,485,760) to the NTT
 * here.  Both compute the exact integer product, so the result is identical; the FFT's own
 * rounding behaviour is not part of this path.  SSA (> NTT max) is rejected with an error.
 *
 * How it is pinned (tests/test_oracle.py): Montgomery against Python pow(), roots against
 * their defining orders, the transform against an O(n^2) DFT, Garner against Python CRT,
 * products against Python int multiplication and closed forms, mul_digits against a
 * Python model of its precision rules.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

#define NP 3
#define NTT_DIGIT_BITS 28
#define DIGIT_MASK ((1u << NTT_DIGIT_BITS) - 1u)
#define MAX_LEN_EXP 26

static const uint32_t P[NP] = {469762049u, 2013265921u, 2281701377u};
static const uint32_t G[NP] = {3u, 31u, 3u};

static uint32_t neg_inv_P[NP], R_sq[NP], vec_one[NP], inv_G[NP];
static uint32_t steps[2][MAX_LEN_EXP][NP];
static uint32_t inv_pow2[MAX_LEN_EXP][NP];
static int inited = 0;

/* bigint_base.h:110-122 */
static int64_t modinv_i64(int64_t a, int64_t m) {
    int64_t x1 = 1, x2 = 0, x3 = 0, x4 = 1, b = m;
    while (b != 0) {
        int64_t c = a / b;
        int64_t n1 = x3, n2 = x4, n3 = x1 - x3 * c, n4 = x2 - x4 * c, na = b, nb = a - b * c;
        x1 = n1; x2 = n2; x3 = n3; x4 = n4; a = na; b = nb;
    }
    return ((x1 % m) + m) % m;
}

/* bigint_base.h:71-80 */
static uint32_t powmod(uint32_t base, uint32_t e, uint32_t mod) {
    uint64_t r = 1, b = base;
    while (e) {
        if (e & 1) r = r * b % mod;
        b = b * b % mod;
        e >>= 1;
    }
    return (uint32_t)r;
}

/* ntt_base.h:65-71: REDC with w = 32 */
uint32_t ora_mont_mul(int I, uint32_t a, uint32_t b) {
    uint64_t t = (uint64_t)a * b;
    uint32_t u = (uint32_t)t * neg_inv_P[I];
    uint64_t t2 = t + (uint64_t)u * P[I];
    uint32_t r = (uint32_t)(t2 >> 32);
    return r < P[I] ? r : r - P[I];
}

static uint32_t add_mod(uint32_t a, uint32_t b, uint32_t p) { /* bigint_base.h:93-99 */
    uint32_t d = p - b;
    return a < d ? a + b : a - d;
}
static uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t p) { /* bigint_base.h:102-108 */
    uint32_t t = a - b;
    return t + ((0u - (uint32_t)(t > a)) & p);
}

static void vec_mul(uint32_t r[NP], const uint32_t a[NP], const uint32_t b[NP]) {
    for (int I = 0; I < NP; ++I) r[I] = ora_mont_mul(I, a[I], b[I]);
}

void ora_init(void) {
    if (inited) return;
    for (int I = 0; I < NP; ++I) {
        int64_t inv = modinv_i64(P[I], (int64_t)1 << 32);          /* ntt_base.h:45-52 */
        neg_inv_P[I] = (uint32_t)(((uint64_t)1 << 32) - (uint64_t)inv);
        uint64_t rmod = ((uint64_t)1 << 32) % P[I];                  /* ntt_base.h:55-62 */
        R_sq[I] = (uint32_t)(rmod * rmod % P[I]);
        vec_one[I] = (uint32_t)rmod;                                 /* ntt_base.h:63 */
        inv_G[I] = (uint32_t)modinv_i64(G[I], P[I]);                 /* ntt_base.h:29-35 */
    }
    for (int I = 0; I < NP; ++I) {                                   /* ntt_base.h:195-208 */
        for (int level = 1; level <= MAX_LEN_EXP; ++level) {
            steps[0][level - 1][I] = powmod(G[I], (P[I] - 1) >> level, P[I]);
            steps[1][level - 1][I] = powmod(inv_G[I], (P[I] - 1) >> level, P[I]);
        }
    }
    for (int level = 0; level < MAX_LEN_EXP; ++level) {
        for (int I = 0; I < NP; ++I) {
            steps[0][level][I] = ora_mont_mul(I, steps[0][level][I], R_sq[I]);
            steps[1][level][I] = ora_mont_mul(I, steps[1][level][I], R_sq[I]);
        }
    }
    for (int I = 0; I < NP; ++I) {                                   /* ntt_base.h:210-222 */
        inv_pow2[0][I] = (uint32_t)modinv_i64(2, P[I]);
        for (int i = 1; i < MAX_LEN_EXP; ++i)
            inv_pow2[i][I] = (uint32_t)((uint64_t)inv_pow2[i - 1][I] * inv_pow2[0][I] % P[I]);
    }
    for (int i = 0; i < MAX_LEN_EXP; ++i)
        for (int I = 0; I < NP; ++I) inv_pow2[i][I] = ora_mont_mul(I, inv_pow2[i][I], R_sq[I]);
    inited = 1;
}

/* Exported constant tables, for the constant-pinning tests. */
void ora_get_constants(uint32_t* p, uint32_t* g, uint32_t* nip, uint32_t* rsq, uint32_t* one,
                       uint32_t* steps_out /* [2][26][3] */, uint32_t* inv2_out /* [26][3] */) {
    ora_init();
    memcpy(p, P, sizeof P);
    memcpy(g, G, sizeof G);
    memcpy(nip, neg_inv_P, sizeof neg_inv_P);
    memcpy(rsq, R_sq, sizeof R_sq);
    memcpy(one, vec_one, sizeof vec_one);
    memcpy(steps_out, steps, sizeof steps);
    memcpy(inv2_out, inv_pow2, sizeof inv_pow2);
}

/* ntt_base.h:228-246; the 128-bit value is returned as lo/hi 64-bit words */
void ora_garner(uint32_t c0, uint32_t c1, uint32_t c2, uint64_t* lo, uint64_t* hi) {
    const uint32_t inv01 = (uint32_t)modinv_i64(P[0], P[1]);
    const uint32_t inv02 = (uint32_t)modinv_i64(P[0], P[2]);
    const uint32_t inv12 = (uint32_t)modinv_i64(P[1], P[2]);
    uint32_t v0 = c0;
    uint32_t v1 = (uint32_t)((uint64_t)sub_mod(c1, v0, P[1]) * inv01 % P[1]);
    uint32_t v2 = (uint32_t)((uint64_t)sub_mod(c2, v0, P[2]) * inv02 % P[2]);
    v2 = (uint32_t)((uint64_t)sub_mod(v2, v1, P[2]) * inv12 % P[2]);
    u128 x = (u128)((uint64_t)P[0] * P[1]);
    x *= v2;
    x += (u128)((uint64_t)P[0] * v1);
    x += v0;
    *lo = (uint64_t)x;
    *hi = (uint64_t)(x >> 64);
}

/* ntt_base.h:268-287: bit reversal applied to each residue class mod NP */
static void multi_bit_swap(uint32_t* v, size_t n) {
    for (size_t i = 1, j = n >> 1; i < n - 1; ++i) {
        if (i < j) {
            for (int u = 0; u < NP; ++u) {
                uint32_t t = v[NP * i + u];
                v[NP * i + u] = v[NP * j + u];
                v[NP * j + u] = t;
            }
        }
        size_t k = n >> 1;
        while (j >= k) { j -= k; k >>= 1; }
        j += k;
    }
}

/* ntt_multithread.cpp:292-309 */
static void imm_mul_num(uint32_t* v, size_t n3, const uint32_t nums[NP]) {
    for (size_t i = 0; i < n3; i += NP)
        for (int I = 0; I < NP; ++I) v[i + I] = ora_mont_mul(I, v[i + I], nums[I]);
}

/* ntt_multithread.cpp:311-328 */
static void imm_mul_vec(uint32_t* v1, const uint32_t* v2, size_t n3) {
    for (size_t i = 0; i < n3; i += NP)
        for (int I = 0; I < NP; ++I) v1[i + I] = ora_mont_mul(I, v1[i + I], v2[i + I]);
}

static int ctz64(uint64_t x) { return __builtin_ctzll(x); }

/* ntt_multithread.cpp:333-379 (serial branch 344-369, which the threaded branch equals) */
void ora_imm_ntt(uint32_t* v, size_t n, int rev) {
    ora_init();
    if (n == 1) return;
    multi_bit_swap(v, n);
    for (size_t len = 2; len <= n; len *= 2) {
        size_t half = len / 2;
        const uint32_t* step = steps[rev][ctz64(half)];
        for (size_t i = 0; i < n; i += len) {
            uint32_t cur[NP], t[NP];
            memcpy(cur, vec_one, sizeof cur);
            for (size_t j = 0; j < half; ++j) {
                size_t ia = (i + j) * NP, ib = (i + j + half) * NP;
                for (int I = 0; I < NP; ++I) {
                    uint32_t a = v[ia + I];
                    uint32_t b = ora_mont_mul(I, cur[I], v[ib + I]);
                    v[ia + I] = add_mod(a, b, P[I]);
                    v[ib + I] = sub_mod(a, b, P[I]);
                }
                vec_mul(t, cur, step);
                memcpy(cur, t, sizeof t);
            }
        }
    }
    if (rev) imm_mul_num(v, n * NP, inv_pow2[ctz64(n) - 1]);
}

/* ntt_multithread.cpp:384-403 */
static void crt_merge(uint32_t* v, size_t n3) {
    for (size_t i = 0; i < n3; i += NP) {
        uint64_t lo, hi;
        ora_garner(v[i], v[i + 1], v[i + 2], &lo, &hi);
        u128 x = ((u128)hi << 64) | lo;
        v[i + 0] = (uint32_t)(x & DIGIT_MASK);
        x >>= NTT_DIGIT_BITS;
        v[i + 1] = (uint32_t)(x & DIGIT_MASK);
        x >>= NTT_DIGIT_BITS;
        v[i + 2] = (uint32_t)(x & DIGIT_MASK);
    }
}

/* ntt_multithread.cpp:418-442; v has capacity >= 3n; returns the trimmed limb count */
static size_t trim(uint32_t* v, size_t n) {
    size_t n3 = n * NP;
    if (n != 1) {
        uint32_t digit = v[1] + v[3];
        v[1] = digit & DIGIT_MASK;
        uint32_t carry = digit >> NTT_DIGIT_BITS;
        for (size_t i = 2, j = i * NP; i < n; ++i, j += NP) {
            digit = v[j - 4] + v[j - 2] + v[j] + carry;
            v[i] = digit & DIGIT_MASK;
            carry = digit >> NTT_DIGIT_BITS;
        }
        digit = v[n3 - 4] + v[n3 - 2] + carry;
        v[n] = digit & DIGIT_MASK;
        carry = digit >> NTT_DIGIT_BITS;
        digit = v[n3 + 3 - 4] + carry;
        v[n + 1] = digit & DIGIT_MASK;
        carry = digit >> NTT_DIGIT_BITS;
        v[n + 2] = carry;
    }
    size_t i = n == 1 ? 2 : n + 2;
    while (v[i] == 0) --i;
    return i + 1;
}

static size_t ceil_div(size_t a, size_t b) { return a / b + (a % b != 0); }

static size_t bit_ceil(size_t x) {
    size_t r = 1;
    while (r < x) r <<= 1;
    return r;
}

/* ntt_multithread.cpp:447-475 with digit_bits = 28, k = NP; returns count written */
static size_t vec64_to_vec32_and_repeat(const uint64_t* src, size_t n, uint32_t* dst) {
    const uint32_t mask = DIGIT_MASK;
    u128 tmp = 0;
    int tmp_bits = 0;
    size_t i = 0, w = 0;
    while (i < n) {
        if (tmp_bits < NTT_DIGIT_BITS && i < n) {
            tmp |= (u128)src[i] << tmp_bits;
            tmp_bits += 64;
            ++i;
        }
        while (tmp_bits >= NTT_DIGIT_BITS) {
            for (int j = 0; j < NP; ++j) dst[w++] = (uint32_t)(tmp & mask);
            tmp >>= NTT_DIGIT_BITS;
            tmp_bits -= NTT_DIGIT_BITS;
        }
    }
    while (tmp) {
        for (int j = 0; j < NP; ++j) dst[w++] = (uint32_t)(tmp & mask);
        tmp >>= NTT_DIGIT_BITS;
    }
    return w;
}

/*
 * ntt_multithread.cpp:444-530.  out must hold na + nb + 1 words.
 * same != 0 selects the squaring branch (A.data() == B.data()).  Returns 0, or -1 on OOM.
 */
int ora_ntt_mul_u64(const uint64_t* A, size_t na, const uint64_t* B, size_t nb, int same,
                    uint64_t* out, size_t* nout) {
    ora_init();
    size_t n = bit_ceil(ceil_div(na * 64, NTT_DIGIT_BITS) + ceil_div(nb * 64, NTT_DIGIT_BITS) - 1);
    size_t n3 = n * NP;
    /* trim writes up to index n + 2 and the repeat writer may emit a few digits past n */
    size_t cap = n3 + 3 * NP + 8;
    uint32_t* va = (uint32_t*)calloc(cap, sizeof(uint32_t));
    if (!va) return -1;
    vec64_to_vec32_and_repeat(A, na, va);
    imm_mul_num(va, n3, R_sq);
    ora_imm_ntt(va, n, 0);
    if (same) {
        imm_mul_vec(va, va, n3);
    } else {
        uint32_t* vb = (uint32_t*)calloc(cap, sizeof(uint32_t));
        if (!vb) { free(va); return -1; }
        vec64_to_vec32_and_repeat(B, nb, vb);
        imm_mul_num(vb, n3, R_sq);
        ora_imm_ntt(vb, n, 0);
        imm_mul_vec(va, vb, n3);
        free(vb);
    }
    ora_imm_ntt(va, n, 1);
    const uint32_t ones[NP] = {1, 1, 1};
    imm_mul_num(va, n3, ones);
    crt_merge(va, n3);
    size_t len = trim(va, n);
    /* repack 28-bit -> u64 (530 lines 504-527) */
    size_t w = 0;
    u128 tmp = 0;
    int tmp_bits = 0;
    size_t i = 0;
    while (i < len) {
        while (tmp_bits < 64 && i < len) {
            tmp |= (u128)va[i] << tmp_bits;
            tmp_bits += NTT_DIGIT_BITS;
            ++i;
        }
        if (tmp_bits >= 64) {
            out[w++] = (uint64_t)tmp;
            tmp_bits -= 64;
            tmp >>= 64;
        }
    }
    if (tmp) {
        out[w++] = (uint64_t)tmp;
    } else {
        while (w > 0 && out[w - 1] == 0) --w;
    }
    *nout = w;
    free(va);
    return 0;
}

/* mul.h:82-102 */
void ora_brute_mul_u64(const uint64_t* A, size_t n, const uint64_t* B, size_t m, uint64_t* res,
                       size_t* nout) {
    memset(res, 0, (n + m) * sizeof(uint64_t));
    for (size_t i = 0; i < n; ++i) {
        uint64_t carry = 0;
        size_t j = 0;
        for (; j < m; ++j) {
            u128 t = (u128)A[i] * B[j] + res[i + j] + carry;
            res[i + j] = (uint64_t)t;
            carry = (uint64_t)(t >> 64);
        }
        for (; carry; ++j) {
            u128 t = (u128)res[i + j] + carry;
            res[i + j] = (uint64_t)t;
            carry = (uint64_t)(t >> 64);
        }
    }
    size_t w = n + m;
    while (w > 0 && res[w - 1] == 0) --w;
    *nout = w;
}

#define BRUTE_MAX_TOTAL_BITS ((size_t)2 * 128 * 64)                       /* mul.h:80 */
#define NTT_MAX_TOTAL_BITS (((size_t)(1u << MAX_LEN_EXP) - 2) * NTT_DIGIT_BITS) /* mul.h:192 */

/* mul.h:208-220 (FFT band routed to the NTT; see header) */
int ora_mul_u64(const uint64_t* A, size_t na, const uint64_t* B, size_t nb, int same,
                uint64_t* out, size_t* nout) {
    size_t total_bits = (na + nb) * 64;
    if (total_bits <= BRUTE_MAX_TOTAL_BITS) {
        ora_brute_mul_u64(A, na, B, nb, out, nout);
        return 0;
    }
    if (total_bits <= NTT_MAX_TOTAL_BITS) return ora_ntt_mul_u64(A, na, B, nb, same, out, nout);
    return -2;
}

/* mul.h:245-268 */
static size_t digits_to_vec64(const uint32_t* x, size_t n, size_t offset, uint64_t* res) {
    u128 tmp = 0;
    size_t tmp_bits = 0, i = offset, w = 0;
    while (i < n) {
        while (tmp_bits < 64 && i < n) {
            tmp |= (u128)x[i] << tmp_bits;
            tmp_bits += NTT_DIGIT_BITS;
            ++i;
        }
        res[w++] = (uint64_t)tmp;
        tmp >>= 64;
        tmp_bits -= 64;
    }
    if (tmp) {
        res[w++] = (uint64_t)tmp;
    } else {
        while (w > 0 && res[w - 1] == 0) --w;
    }
    return w;
}

/*
 * mul.h:223-318.  Both inputs must be nonzero (the reference's precondition).
 * out must hold na + nb + 4 limbs.  *point_delta receives the amount subtracted from the
 * caller's point position (offset_a + offset_b + offset_res) when precision != 0, else 0.
 * Returns 0, -1 on OOM, -2 when the product is beyond the NTT's range.
 */
int ora_mul_digits(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, int same,
                   size_t precision, uint32_t* out, size_t* nout, int64_t* point_delta) {
    size_t offset_a = precision == 0 || na <= precision ? 0 : na - precision;
    size_t offset_b = precision == 0 || nb <= precision ? 0 : nb - precision;
    size_t tz_a = 0, tz_b = 0;
    while (a[offset_a + tz_a] == 0) ++tz_a;
    if (same) {
        tz_b = tz_a;
    } else {
        while (b[offset_b + tz_b] == 0) ++tz_b;
    }
    offset_a += tz_a;
    offset_b += tz_b;
    uint64_t* A = (uint64_t*)malloc((ceil_div((na - offset_a) * NTT_DIGIT_BITS, 64) + 2) * 8);
    uint64_t* B = same ? A : (uint64_t*)malloc((ceil_div((nb - offset_b) * NTT_DIGIT_BITS, 64) + 2) * 8);
    if (!A || !B) { free(A); if (!same) free(B); return -1; }
    size_t wa = digits_to_vec64(a, na, offset_a, A);
    size_t wb = same ? wa : digits_to_vec64(b, nb, offset_b, B);
    uint64_t* C = (uint64_t*)malloc((wa + wb + 2) * 8);
    if (!C) { free(A); if (!same) free(B); return -1; }
    size_t wc = 0;
    int rc = ora_mul_u64(A, wa, B, wb, same, C, &wc);
    free(A);
    if (!same) free(B);
    if (rc) { free(C); return rc; }
    /* mul.h:274-300 */
    size_t r = 0;
    {
        u128 tmp = 0;
        size_t tmp_bits = 0, i = 0;
        while (i < wc) {
            if (tmp_bits < NTT_DIGIT_BITS) {
                tmp |= (u128)C[i] << tmp_bits;
                tmp_bits += 64;
                ++i;
            }
            while (tmp_bits >= NTT_DIGIT_BITS) {
                out[r++] = (uint32_t)(tmp & DIGIT_MASK);
                tmp >>= NTT_DIGIT_BITS;
                tmp_bits -= NTT_DIGIT_BITS;
            }
        }
        if (tmp) {
            while (tmp) {
                out[r++] = (uint32_t)(tmp & DIGIT_MASK);
                tmp >>= NTT_DIGIT_BITS;
            }
        } else {
            while (r > 0 && out[r - 1] == 0) --r;
        }
    }
    free(C);
    /* mul.h:302-315 */
    if (precision) {
        size_t offset_res = r <= precision ? 0 : r - precision;
        while (out[offset_res] == 0) ++offset_res;
        if (offset_res) {
            memmove(out, out + offset_res, (r - offset_res) * sizeof(uint32_t));
            r -= offset_res;
        }
        *point_delta = (int64_t)(offset_a + offset_b + offset_res);
    } else {
        size_t z = tz_a + tz_b;
        if (z) {
            memmove(out + z, out, r * sizeof(uint32_t));
            memset(out, 0, z * sizeof(uint32_t));
            r += z;
        }
        *point_delta = 0;
    }
    *nout = r;
    return 0;
}

/* Convenience for tests: the product of two 28-bit limb arrays via mul_digits, precision 0. */
int ora_mul_limbs(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, int same,
                  uint32_t* out, size_t* nout) {
    int64_t pd;
    return ora_mul_digits(a, na, b, nb, same, 0, out, nout, &pd);
}
