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

// Extremes of the positions and radii for the overlap grid, in one
// workgroup: out = {min x, min y, min z, max x, max y, max z, max r, all
// finite ? 1 : 0}.
__global__ __launch_bounds__(1024) void k_overlap_bounds(const double4* __restrict__ x,
                                                         uint32_t np, double* __restrict__ out) {
    __shared__ double red[8][1024];
    double v[8] = {__builtin_inf(), __builtin_inf(), __builtin_inf(), -__builtin_inf(),
                   -__builtin_inf(), -__builtin_inf(), 0.0, 1.0};
    for (uint32_t i = threadIdx.x; i < np; i += 1024) {
        const double4 p = x[i];
        v[0] = fmin(v[0], p.x);
        v[1] = fmin(v[1], p.y);
        v[2] = fmin(v[2], p.z);
        v[3] = fmax(v[3], p.x);
        v[4] = fmax(v[4], p.y);
        v[5] = fmax(v[5], p.z);
        v[6] = fmax(v[6], p.w);
        if (!(__builtin_isfinite(p.x) && __builtin_isfinite(p.y) && __builtin_isfinite(p.z) &&
              __builtin_isfinite(p.w) && p.w >= 0.0))
            v[7] = 0.0;
    }
    for (int k = 0; k < 8; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (uint32_t h = 512; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            const uint32_t o = threadIdx.x + h;
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][o]);
            for (int k = 3; k < 7; ++k) red[k][threadIdx.x] = fmax(red[k][threadIdx.x], red[k][o]);
            red[7][threadIdx.x] = fmin(red[7][threadIdx.x], red[7][o]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 8) out[threadIdx.x] = red[threadIdx.x][0];
}

__device__ __forceinline__ int grid_coord(double v, double lo, double inv_h, int n) {
    const double a = (v - lo) * inv_h;
    int c = a >= 0.0 ? (int)fmin(a, (double)(n - 1)) : 0;
    return c;
}

__device__ __forceinline__ uint32_t grid_key(const double4& p, const OverlapGrid& g) {
    const int ix = grid_coord(p.x, g.lo[0], g.inv_h, g.n[0]);
    const int iy = grid_coord(p.y, g.lo[1], g.inv_h, g.n[1]);
    const int iz = grid_coord(p.z, g.lo[2], g.inv_h, g.n[2]);
    return ((uint32_t)iz * (uint32_t)g.n[1] + (uint32_t)iy) * (uint32_t)g.n[0] + (uint32_t)ix;
}

__global__ __launch_bounds__(256) void k_overlap_bin(const double4* __restrict__ x, uint32_t np,
                                                     OverlapGrid g, uint32_t* __restrict__ key,
                                                     uint32_t* __restrict__ cell_count) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= np) return;
    const uint32_t k = grid_key(x[i], g);
    key[i] = k;
    atomicAdd(&cell_count[k], 1u);
}

__global__ __launch_bounds__(256) void k_overlap_scatter(const uint32_t* __restrict__ key,
                                                         uint32_t np,
                                                         uint32_t* __restrict__ cursor,
                                                         uint32_t* __restrict__ order) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= np) return;
    order[atomicAdd(&cursor[key[i]], 1u)] = i;
}

// The overlap test of DEM::findOverlaps (src/sphere.cpp:694-716) for the
// j > i in the 27 cells around i's.  Pass 1 (pairs == nullptr) counts;
// pass 2 writes from offs[i] on and then orders the run by j (insertion
// sort: a particle has a handful of overlaps), which gives the reference's
// i-major, j-ascending list whatever order the cells hold.
__global__ __launch_bounds__(256) void k_overlaps_grid(
    const double4* __restrict__ x, const double4* __restrict__ vel, uint32_t np, OverlapGrid g,
    const uint32_t* __restrict__ cell_start, const uint32_t* __restrict__ order,
    uint32_t* __restrict__ counts, const uint64_t* __restrict__ offs, uint32_t* __restrict__ pairs,
    double* __restrict__ delta) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= np) return;
    const double4 xi = x[i];
    if (vel[i].w != 0) {
        if (!pairs) counts[i] = 0;
        return;
    }
    const int cx = grid_coord(xi.x, g.lo[0], g.inv_h, g.n[0]);
    const int cy = grid_coord(xi.y, g.lo[1], g.inv_h, g.n[1]);
    const int cz = grid_coord(xi.z, g.lo[2], g.inv_h, g.n[2]);
    uint32_t n = 0;
    const uint64_t w0 = pairs ? offs[i] : 0;
    for (int z = cz - 1; z <= cz + 1; ++z) {
        if (z < 0 || z >= g.n[2]) continue;
        for (int y = cy - 1; y <= cy + 1; ++y) {
            if (y < 0 || y >= g.n[1]) continue;
            for (int xx = cx - 1; xx <= cx + 1; ++xx) {
                if (xx < 0 || xx >= g.n[0]) continue;
                const uint32_t c = ((uint32_t)z * (uint32_t)g.n[1] + (uint32_t)y) *
                                       (uint32_t)g.n[0] + (uint32_t)xx;
                for (uint32_t k = cell_start[c]; k < cell_start[c + 1]; ++k) {
                    const uint32_t j = order[k];
                    if (j <= i) continue;
                    const double4 xj = x[j];
                    const double dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                    const double len = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
                    const double dn = len - (xi.w + xj.w);
                    if (dn < 0.0 && vel[j].w == 0) {
                        if (pairs) {
                            pairs[2 * (w0 + n)] = i;
                            pairs[2 * (w0 + n) + 1] = j;
                            delta[w0 + n] = dn;
                        }
                        ++n;
                    }
                }
            }
        }
    }
    if (!pairs) {
        counts[i] = n;
        return;
    }
    for (uint32_t a = 1; a < n; ++a) {
        const uint32_t jj = pairs[2 * (w0 + a) + 1];
        const double dd = delta[w0 + a];
        uint32_t b = a;
        while (b > 0 && pairs[2 * (w0 + b - 1) + 1] > jj) {
            pairs[2 * (w0 + b) + 1] = pairs[2 * (w0 + b - 1) + 1];
            delta[w0 + b] = delta[w0 + b - 1];
            --b;
        }
        pairs[2 * (w0 + b) + 1] = jj;
        delta[w0 + b] = dd;
    }
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

hipError_t launch_overlap_bounds(const double4* x, uint32_t np, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_overlap_bounds, dim3(1), dim3(1024), 0, s, x, np, out);
    return hipGetLastError();
}

hipError_t launch_overlap_bin(const double4* x, uint32_t np, const OverlapGrid& g, uint32_t* key,
                              uint32_t* cell_count, hipStream_t s) {
    if (np == 0) return hipSuccess;
    hipLaunchKernelGGL(k_overlap_bin, dim3((np + 255) / 256), dim3(256), 0, s, x, np, g, key,
                       cell_count);
    return hipGetLastError();
}

hipError_t launch_overlap_scatter(const uint32_t* key, uint32_t np, uint32_t* cursor,
                                  uint32_t* order, hipStream_t s) {
    if (np == 0) return hipSuccess;
    hipLaunchKernelGGL(k_overlap_scatter, dim3((np + 255) / 256), dim3(256), 0, s, key, np, cursor,
                       order);
    return hipGetLastError();
}

hipError_t launch_overlaps_grid(const double4* x, const double4* vel, uint32_t np,
                                const OverlapGrid& g, const uint32_t* cell_start,
                                const uint32_t* order, uint32_t* counts, const uint64_t* offs,
                                uint32_t* pairs, double* delta, hipStream_t s) {
    if (np == 0) return hipSuccess;
    hipLaunchKernelGGL(k_overlaps_grid, dim3((np + 255) / 256), dim3(256), 0, s, x, vel, np, g,
                       cell_start, order, counts, offs, pairs, delta);
    return hipGetLastError();
}

hipError_t launch_porosity(const double4* x, uint32_t np, int z_slices, double h_slice,
                           double V_slice, double* z_pos, double* porosity, hipStream_t s) {
    hipLaunchKernelGGL(k_porosity, dim3((z_slices + 63) / 64), dim3(64), 0, s, x, np, z_slices,
                       h_slice, V_slice, z_pos, porosity);
    return hipGetLastError();
}

}  // namespace sphx
