// Output-side readers of the step's state on the GPU: the pairwise overlap
// list of DEM::findOverlaps (src/sphere.cpp:674-724, behind `sphere
// --contacts`) and the porosity profile of DEM::porosity
// (src/sphere.cpp:486-592).  Both give the reference's results exactly: the
// same expressions in the same order, FP64 throughout.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "analysis.hpp"

namespace sphx {

namespace {

constexpr uint32_t TILE = 256;
constexpr double PI_D = 3.14159265358979323846;  // M_PI

// Pass 1 (pairs == nullptr): count, per i, the overlapping j > i.  Pass 2:
// write them from offs[i] on, in ascending j.  Workgroup b takes
// i in [256 b, 256 b + 256) and walks the j tiles from its own onwards
// through LDS.
__global__ __launch_bounds__(TILE) void k_overlaps(const double4* __restrict__ x,
                                                   const double4* __restrict__ vel, uint32_t np,
                                                   uint32_t* __restrict__ counts,
                                                   const uint64_t* __restrict__ offs,
                                                   uint32_t* __restrict__ pairs,
                                                   double* __restrict__ delta) {
    __shared__ double4 tx[TILE];
    __shared__ double tfix[TILE];
    const uint32_t i = blockIdx.x * TILE + threadIdx.x;
    const bool in = i < np;
    const double4 xi = in ? x[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    const double fi = in ? vel[i].w : 1.0;
    uint32_t n = 0;
    uint64_t w = (in && pairs) ? offs[i] : 0;
    for (uint32_t t0 = blockIdx.x * TILE; t0 < np; t0 += TILE) {
        const uint32_t j = t0 + threadIdx.x;
        if (j < np) {
            tx[threadIdx.x] = x[j];
            tfix[threadIdx.x] = vel[j].w;
        }
        __syncthreads();
        const uint32_t m = np - t0 < TILE ? np - t0 : TILE;
        if (in && fi == 0) {
            for (uint32_t k = 0; k < m; ++k) {
                const uint32_t jj = t0 + k;
                if (jj <= i) continue;
                const double4 xj = tx[k];
                const double dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                const double len = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
                const double dn = len - (xi.w + xj.w);
                if (dn < 0.0 && tfix[k] == 0) {
                    if (pairs) {
                        pairs[2 * w] = i;
                        pairs[2 * w + 1] = jj;
                        delta[w] = dn;
                        ++w;
                    } else {
                        ++n;
                    }
                }
            }
        }
        __syncthreads();
    }
    if (in && !pairs) counts[i] = n;
}

__device__ __forceinline__ double spherical_cap(double h, double r) {
    return PI_D * h * h / 3.0 * (3.0 * r - h);
}

// One work-item per slice, the particles in index order (the reference's
// serial subtraction per slice).
__global__ __launch_bounds__(64) void k_porosity(const double4* __restrict__ x, uint32_t np,
                                                 int z_slices, double h_slice, double V_slice,
                                                 double* __restrict__ z_pos,
                                                 double* __restrict__ porosity) {
    const int iz = (int)(blockIdx.x * 64 + threadIdx.x);
    if (iz >= z_slices) return;
    double V_void = V_slice;
    const double lo = iz * h_slice;
    const double hi = lo + h_slice;
    for (uint32_t i = 0; i < np; ++i) {
        const double4 xi = x[i];
        const double zc = xi.z, r = xi.w;
        const double zl = zc - r, zh = zc + r;
        const double V_sphere = 4.0 / 3.0 * PI_D * r * r * r;
        if (lo <= zl && zh <= hi) {
            V_void -= V_sphere;
        } else {
            if (lo >= zc && lo <= zh)
                V_void -= spherical_cap(zh - lo, r);
            else if (lo < zc && lo > zl)
                V_void -= V_sphere + spherical_cap(lo - zl, r);
            if (hi <= zc && hi >= zl)
                V_void -= spherical_cap(hi - zl, r);
            else if (hi > zc && hi < zh)
                V_void -= V_sphere + spherical_cap(zh - hi, r);
        }
    }
    z_pos[iz] = lo + 0.5 * h_slice;
    porosity[iz] = V_void / V_slice;
}

}  // namespace

hipError_t launch_overlaps(const double4* x, const double4* vel, uint32_t np, uint32_t* counts,
                           const uint64_t* offs, uint32_t* pairs, double* delta, hipStream_t s) {
    if (np == 0) return hipSuccess;
    hipLaunchKernelGGL(k_overlaps, dim3((np + TILE - 1) / TILE), dim3(TILE), 0, s, x, vel, np,
                       counts, offs, pairs, delta);
    return hipGetLastError();
}

hipError_t launch_porosity(const double4* x, uint32_t np, int z_slices, double h_slice,
                           double V_slice, double* z_pos, double* porosity, hipStream_t s) {
    hipLaunchKernelGGL(k_porosity, dim3((z_slices + 63) / 64), dim3(64), 0, s, x, np, z_slices,
                       h_slice, V_slice, z_pos, porosity);
    return hipGetLastError();
}

}  // namespace sphx
