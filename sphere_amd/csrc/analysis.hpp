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

// DEM::porosity of z_slices slices of height h_slice (volume V_slice).
hipError_t launch_porosity(const double4* x, uint32_t np, int z_slices, double h_slice,
                           double V_slice, double* z_pos, double* porosity, hipStream_t s);

}  // namespace sphx

#endif
