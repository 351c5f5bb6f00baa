// Argument blocks shared by the kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bi {

enum PassFlags : uint32_t {
    kFromLimbs = 1u,  // the pass reads the operands' 32-bit words from their limbs
    kSquare = 2u,     // B is A: one forward transform, pointwise square
    kSingle = 4u,     // strided forward passes transform operand A only
    kBPre = 8u,       // block pass: res_b already holds B's full forward transform
    kFwdOnly = 16u,   // block pass: forward-transform A and store it (no product, no inverse)
    kInvOnly = 32u,   // block pass: res_a holds a product spectrum; inverse stages only
};

// Transform inputs are 32-bit words of the operands' bit strings (word g = bits [32g, 32g+32)),
// cut from the 28-bit limbs on load: 8 limbs make 7 words.
constexpr uint32_t words_of(uint32_t limbs) { return (uint32_t)(((uint64_t)limbs * 7u + 7u) / 8u); }

// One product inside a batched launch.  Residue buffers are planar: 3 planes of n u32.
struct PassJob {
    const uint32_t* limbs_a;
    const uint32_t* limbs_b;
    uint32_t la, lb;      // limb counts of the whole operands
    uint32_t* res_a;
    uint32_t* res_b;
    uint32_t a_off;       // A's transform takes words [a_off, a_off + a_cnt) (a chunk of A)
    uint32_t a_cnt;       // ~0u: all of A
};

constexpr uint32_t kBetaPowStride = 16;  // words per block in the beta power tables

struct PassArgs {
    PassJob job;            // used when jobs == nullptr
    const PassJob* jobs;    // device array, indexed by blockIdx.y
    uint32_t log_n;         // transform length n = 2^log_n
    uint32_t log_r, log_c;  // tile: R rows x C columns
    uint32_t log_s;         // row stride S
    uint32_t flags;
    const uint32_t* wtab;      // forward root table: [prime][4096], Montgomery form
    // per-block roots of this pass: beta^(2^e) for e < log_r, [prime][beta_stride][kBetaPowStride],
    // or null; forward and inverse stages use the same roots
    const uint32_t* beta;
    uint32_t beta_stride;      // blocks per prime
    // The first kernel of a launch sequence zeroes these words (the carry-scan status words and
    // fix-up flags of the sequence's Garner launch) before anything else; null elsewhere.
    uint32_t* clear;
    uint32_t clear_words;
};

struct GarnerConsts {
    uint32_t inv01, inv02, inv12;  // P0^-1 mod P1, P0^-1 mod P2, P1^-1 mod P2 (Montgomery form)
};

// Output tiles of the Garner + carry kernel: 4096 limbs = 3584 words (114688 bits).
constexpr uint32_t kCarryTileLimbs = 4096;
constexpr uint32_t kCarryTileWords = 3584;

struct GarnerJob {
    const uint32_t* y;   // 3 planes of n: inverse transform output (times n * R^-1), coefficient j
                         // at position (n - j) mod n
    uint32_t* out;       // out_len limbs
    uint32_t n, out_len;
    uint32_t kscale[3];  // n^-1 * R^2 mod P_i, so mont_mul(y, kscale) = y * n^-1 * R
    uint32_t* status;    // one word per output tile: its carry map, and kFixBit when the tile
                         // was written with an assumed carry-in of 0
    uint32_t* fix;       // set to 1 by any such tile (zeroed before launch): the fix-up runs
    // Chunked product: the long operand was cut into nchunks pieces of chunk_len words, each
    // multiplied with a transform of length n >= chunk_len + (words of B) - 1; chunk c's
    // residues start at y + c * chunk_stride and its product lands at word offset c * chunk_len.
    // chunk_len is a power of two (log_chunk < 32) or a multiple of kCarryTileWords
    // (log_chunk == 0xFF).  nchunks <= 1 means an ordinary product.
    uint32_t nchunks, log_chunk, chunk_len;
    uint64_t chunk_stride;
    // Unchunked products may use n = (words of A) + (words of B) - 2: the one coefficient past
    // the transform, (top word of A) * (top word of B) at index n, wraps onto index 0.  Garner
    // takes it back out of coefficient 0 and emits it at index n.  wrap_a == null: no wrap.
    const uint32_t* wrap_a;
    const uint32_t* wrap_b;
    uint32_t wrap_la, wrap_lb;
};

struct GarnerArgs {
    GarnerJob job;
    const GarnerJob* jobs;
    GarnerConsts k;
};

hipError_t launch_outer_pass(const PassArgs& a, bool inv, uint32_t njobs, uint32_t nops,
                             hipStream_t st);
hipError_t launch_mid_pass(const PassArgs& a, uint32_t njobs, hipStream_t st);
hipError_t launch_garner_carry(const GarnerArgs& a, uint32_t max_out_len, uint32_t njobs,
                               bool chunked, hipStream_t st);
// Diagonal sums of chunk spectra: diag[s] = sum over i + k = s of spec_a[i] * spec_b[k]
// (pointwise Montgomery products mod each prime), s = 0 .. ca + cb - 2.  Spectra and diagonals
// are 3 planes of n words each, consecutive.
hipError_t launch_diag_accumulate(const uint32_t* spec_a, const uint32_t* spec_b, uint32_t ca,
                                  uint32_t cb, uint32_t n, uint32_t* diag, hipStream_t st);

}  // namespace bi
