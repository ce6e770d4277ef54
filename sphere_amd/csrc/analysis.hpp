// Output-side readers (overlap list, porosity profile) on the GPU.
#ifndef SPHERE_AMD_ANALYSIS_HPP_
#define SPHERE_AMD_ANALYSIS_HPP_

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sphx {

// DEM::findOverlaps over one simulation's np particles (original order):
// with pairs == nullptr, counts[i] = overlapping j > i; else the pairs and
// overlaps written from offs[i] (exclusive prefix of counts).
hipError_t launch_overlaps(const double4* x, const double4* vel, uint32_t np, uint32_t* counts,
                           const uint64_t* offs, uint32_t* pairs, double* delta, hipStream_t s);

// Uniform grid for the overlap search: cells of side 1/inv_h >= the largest
// contact distance, n[d] cells per axis from lo[d]; key = (z n1 + y) n0 + x.
struct OverlapGrid {
    double lo[3];
    double inv_h;
    int n[3];
    int pad;
};

// Position/radius extremes (8 doubles: min xyz, max xyz, max r, finite flag).
hipError_t launch_overlap_bounds(const double4* x, uint32_t np, double* out, hipStream_t s);
// key[i] = cell of particle i; cell_count[key] += 1 (cell_count zeroed).
hipError_t launch_overlap_bin(const double4* x, uint32_t np, const OverlapGrid& g, uint32_t* key,
                              uint32_t* cell_count, hipStream_t s);
// order[cursor[key[i]]++] = i (cursor = exclusive prefix of the counts).
hipError_t launch_overlap_scatter(const uint32_t* key, uint32_t np, uint32_t* cursor,
                                  uint32_t* order, hipStream_t s);
// launch_overlaps through the grid: same counts, same list.
hipError_t launch_overlaps_grid(const double4* x, const double4* vel, uint32_t np,
                                const OverlapGrid& g, const uint32_t* cell_start,
                                const uint32_t* order, uint32_t* counts, const uint64_t* offs,
                                uint32_t* pairs, double* delta, hipStream_t s);

// DEM::porosity of z_slices slices of height h_slice (volume V_slice).
hipError_t launch_porosity(const double4* x, uint32_t np, int z_slices, double h_slice,
                           double V_slice, double* z_pos, double* porosity, hipStream_t s);

}  // namespace sphx

#endif
