// Spatial hash of the SDF obstacles for contact detection (implemented in contact_hash.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "kernels.h"

namespace sbs {

// Cell coordinate floor(x / h) (x * inv_h in FP64), clamped to +-2^30.
__host__ __device__ inline int hash_cell(double x, double inv_h)
{
    double const c = floor(x * inv_h);
    return int(c < -1073741824.0 ? -1073741824.0 : (c > 1073741824.0 ? 1073741824.0 : c));
}

// Slot of cell (i, j, k) in a table of mask + 1 (a power of two) entries (Teschner et al. 2003).
__host__ __device__ inline uint32_t hash_slot(int i, int j, int k, uint32_t mask)
{
    return ((uint32_t(i) * 73856093u) ^ (uint32_t(j) * 19349663u) ^ (uint32_t(k) * 83492791u)) & mask;
}

// The built hash as the detection kernel reads it: slot s holds the obstacle indices
// entries[start[s] .. end[s]) in increasing order; global[0 .. nglobal) (increasing) are the
// obstacles every vertex tests (those not hashed).
struct contact_hash_dev {
    const uint32_t* start;
    const uint32_t* end;
    const uint32_t* entries;
    const uint32_t* global;
    uint32_t nglobal;
    uint32_t mask;
    double inv_h;
};

// Workspace (uint32 words) of build_contact_hash.
std::size_t contact_hash_workspace(uint32_t nhashed, uint64_t pairs);

// Builds the table (2^bits slots) of the obstacles sdfs[hashed_ids[0 .. nhashed)] whose volumes
// cover `pairs` cells in total: start/end get 2^bits words, entries up to `pairs` words, and
// *num_entries (device word) the number of entries after repeated pairs are dropped.
void build_contact_hash(hipStream_t s, const sdf_dev* sdfs, const uint32_t* hashed_ids, uint32_t nhashed,
                        uint64_t pairs, double inv_h, int bits, uint32_t* start, uint32_t* end, uint32_t* entries,
                        uint32_t* num_entries, uint32_t* ws);

void launch_detect_hashed(hipStream_t s, contact_dev c, const pos4* x, const sdf_dev* sdfs, contact_hash_dev h,
                          int block);

}  // namespace sbs
