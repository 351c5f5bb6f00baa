// Batched signed additions on device-resident limb arrays (add_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace bi {

constexpr uint32_t kAddTileLimbs = 4096;

// One job: |X| + |Y| (subtract == 0) or the larger magnitude minus the smaller.  width limbs are
// written: max(lx, ly) + 1 for an add, max(lx, ly) for a subtraction.  maps and carry_in hold
// ntiles = ceil(width / 4096) words each; flags[0] = |Y| > |X| (subtraction), flags[1] = result
// length without leading zeros (0 for zero).
struct AddJob {
    const uint32_t* x;
    size_t lx;
    const uint32_t* y;
    size_t ly;
    uint32_t* out;
    size_t width;
    uint32_t* maps;
    uint32_t* carry_in;
    uint32_t* flags;
    uint32_t subtract;
    uint32_t ntiles;
};

// Three launches on st: tile maps, per-job scan, apply.  d_jobs is a device array.
hipError_t launch_add_batch(const AddJob* d_jobs, uint32_t njobs, uint32_t max_tiles, hipStream_t st);

}  // namespace bi
