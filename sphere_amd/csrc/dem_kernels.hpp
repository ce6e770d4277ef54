// Launch wrappers of the DEM step kernels (definitions in dem_kernels.hip).
#ifndef SPHERE_AMD_DEM_KERNELS_HPP_
#define SPHERE_AMD_DEM_KERNELS_HPP_

#include <hip/hip_runtime.h>

#include "dem_arrays.hpp"
#include "dem_types.hpp"

namespace sphx {

constexpr int RADIX_ITEMS = 8;                        // keys per thread per tile
constexpr int RADIX_TILE = BLOCK * RADIX_ITEMS;       // keys per tile

__host__ __device__ inline unsigned n_tiles(unsigned n) {
    return (n + RADIX_TILE - 1) / RADIX_TILE;
}
__host__ __device__ inline unsigned n_blocks(unsigned n) { return (n + BLOCK - 1) / BLOCK; }

// All launches cover every simulation of the context (DevArrays holds the
// batch layout); grids are whole 256-particle chunks.

// Cell key of every particle (calcParticleCellID, src/sorting.cuh:25-51)
// into the current step's flag slot.
hipError_t launch_hash(const DevArrays& a, hipStream_t s);
// Stable LSD radix sort of (key, identity) into (skey, perm), `passes` digits
// of 8 bits (replaces thrust::sort_by_key, src/device.cu:1075-1078), one
// launch per stage, over the whole batch (keys are global).
hipError_t launch_sort(const DevArrays& a, int passes, int bits, hipStream_t s);
// cellStart reset, cellStart/cellEnd, sorted copies
// (src/device.cu:1089-1090, src/sorting.cuh:56-132).
hipError_t launch_cells(const DevArrays& a, int sorted_bits, hipStream_t s);
// topology + interact fused per particle (src/contactsearch.cuh:194-726);
// `lanes` (1, 4, 8 or 16) work-items per particle for contact models 2 and 3.
hipError_t launch_contact(const DevArrays& a, int contactmodel, int lanes, int wg,
                          hipStream_t s);
// bondsLinear per bond, then a per-particle gather in bond order
// (src/cohesion.cuh:8-233).
// Contact pass (models 2 and 3, four lanes per particle) that also
// integrates each particle and hashes it for the next step; then the wall
// force partials.  For a fixed grid height, without bonds or a wmode-3 wall.
hipError_t launch_contact_integrate(const DevArrays& a, int contactmodel, hipStream_t s);
hipError_t launch_wall_partials(const DevArrays& a, hipStream_t s);
hipError_t launch_bonds(const DevArrays& a, hipStream_t s);
// findShearStressOnFixedMovingParticles + summation + serial sum
// (src/integration.cuh:457-484,309-341,64-68), wmode 3 only.
hipError_t launch_shear_stress(const DevArrays& a, hipStream_t s);
// integrate (TY3), src/integration.cuh:14-305, the wall force block
// partials, and (hash_next != 0) the next step's cell keys.
hipError_t launch_integrate(const DevArrays& a, int hash_next, hipStream_t s);
// Four work-items per particle (one vector component each), hashing for the
// next step; the wall partials then need launch_wall_partials.  Fixed grid
// height, no wmode-3 wall.
hipError_t launch_integrate_q(const DevArrays& a, hipStream_t s);
// Per simulation: serial sum of the partials + integrateWalls + loop advance
// (src/integration.cuh:344-450, src/device.cu:2500-2538).  With resort != 0
// it also re-sorts the simulation for the next step, in one workgroup, when
// the integration pass (launched with hash_next) found a cell change.
hipError_t launch_check_sqrt32(unsigned long long* bad, uint64_t lo, uint64_t hi, hipStream_t s);
hipError_t launch_walls(const DevArrays& a, int resort, int passes, int bits, int sorted_bits,
                        hipStream_t s);
// Fill n words with v.
hipError_t launch_fill_u32(uint32_t* dst, unsigned n, uint32_t v, hipStream_t s);
// n contact-row words indexing the order of perm_rows, rewritten with the
// partners' original indices into out.
hipError_t launch_rows_to_orig(const uint32_t* rows, const uint32_t* perm_rows, uint32_t* out,
                               uint32_t n, hipStream_t s);

}  // namespace sphx

#endif
