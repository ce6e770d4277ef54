// HIP kernels of the DEM particle timestep for CDNA4 (gfx950).
//
// One workgroup = 256 work-items = four 64-wide wavefronts.  All arithmetic
// is FP64 with the reference's operation order; this file must be compiled
// with -ffp-contract=off (no fused multiply-add), as the reference is with
// nvcc --fmad=false (src/CMakeLists.txt:25).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dem_arrays.hpp"
#include "dem_kernels.hpp"
#include "dem_math.hpp"
#include "dem_types.hpp"

namespace sphx {

namespace {

constexpr double PI = 3.14159265358979;  // src/constants.h:14

// The parameter table is read-only for every kernel (the host writes it
// between launches), so it is read through the constant address space:
// uniform fields become scalar loads that stores to global memory do not
// invalidate, instead of vector loads reissued after every store.
typedef const __attribute__((address_space(4))) DevParams CParams;

// The walls' state: only the wall pass writes it, so the kernels before it
// in a step read it through the constant address space as well.
__device__ __forceinline__ double4 const_double4(const double4* v) {
    const auto* c = (const __attribute__((address_space(4))) double*)v;
    return make_double4(c[0], c[1], c[2], c[3]);
}
__device__ __forceinline__ double4 wall_nx(const DevArrays& a, uint32_t w) {
    return const_double4(a.w_nx + w);
}
__device__ __forceinline__ double4 wall_mvfd(const DevArrays& a, uint32_t w) {
    return const_double4(a.w_mvfd + w);
}

__device__ __forceinline__ double grid_Lz(CParams& p, const DevArrays& a) {
    // DEM::updateGridSize (src/device.cu:363-393): with an adaptive grid the
    // step runs with L[2] equal to the top wall position at its start.
    return (p.adaptive == 1 && p.nw > 0) ? wall_nx(a, p.wall_base).w : (double)p.L[2];
}

// calcCellID, src/sorting.cuh:7-20: the integer coordinate is the double
// floor of (x - origo) / (L / num), converted to unsigned.
__device__ __forceinline__ uint32_t to_u32(double v) {
    // Values are >= 0 for particles inside the grid; a negative value (a
    // particle below origo) converts to 0 here, where the reference's
    // conversion is undefined.
    return (uint32_t)v;
}

__device__ __forceinline__ uint32_t calc_cell_id(CParams& p, double Lz, double x,
                                                 double y, double z) {
    const uint32_t ix = to_u32(floor((x - p.origo[0]) / (p.L[0] / p.num[0])));
    const uint32_t iy = to_u32(floor((y - p.origo[1]) / (p.L[1] / p.num[1])));
    const uint32_t iz = to_u32(floor((z - p.origo[2]) / (Lz / p.num[2])));
    return (iz * p.num[1]) * p.num[0] + iy * p.num[0] + ix;
}

// Per-step flags live in slot iter & 1 of the loop state.  All
// simulations of a context step together, so their parities agree.
__device__ __forceinline__ uint32_t step_slot(const DevLoop* lp) {
    return (uint32_t)(lp->iter & 1);
}
// Some simulation of the batch needs this step's re-sort.
__device__ __forceinline__ bool sort_needed(const DevGlobal* g, const DevLoop* lp0) {
    return g->changed[step_slot(lp0)] != 0u;
}

// The set of sorted copies the current step reads (loop-counter parity).
__device__ __forceinline__ uint32_t cur_set(const DevLoop* lp) {
    return (uint32_t)(lp->iter & 1);
}

// The simulation that owns a 256-particle chunk.
__device__ __forceinline__ CParams& sim_params(const DevArrays& a, uint32_t sim) {
    return *(CParams*)(a.ptab + sim);
}

__device__ __forceinline__ CParams& chunk_params(const DevArrays& a, uint32_t chunk) {
    return sim_params(a, a.chunk_sim[chunk]);
}

// Workgroups are dealt to the eight XCDs in turn (workgroup b runs on XCD
// b % 8), and each XCD has its own L2.  With xcd_order set, renumbering
// gives each XCD one contiguous range of workgroups, so that the neighbours
// a workgroup gathers (close in sorted order) tend to sit in the L2 of the
// XCD that reads them.  On the 1e5-particle sheared bed the contact pass
// measured 62.8 us in hardware order against 73.4 us renumbered (the
// renumbered order runs the chip's concurrent workgroups over eight far
// apart ranges), so hardware order is the default.
__device__ __forceinline__ uint32_t xcd_block(const DevArrays& a, uint32_t b, uint32_t n) {
    constexpr uint32_t X = 8;
    if (!a.xcd_order) return b;
    const uint32_t per = n / X, rem = n % X;
    const uint32_t x = b % X, k = b / X;
    return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + k;
}

// ---------------------------------------------------------------------------
// Hash (calcParticleCellID, src/sorting.cuh:25-51), plus the two tests that
// decide whether the step may reuse the previous order and candidate lists.

struct HashFlags {
    uint32_t changed = 0, moved = 0, nonfinite = 0, outside = 0;
};

__device__ __forceinline__ void hash_particle_from(CParams& p, const DevArrays& a, uint32_t i,
                                                   double Lz, const double4 x, uint32_t key_prev,
                                                   const double4 xb, HashFlags& f) {
    uint32_t lk = calc_cell_id(p, Lz, x.x, x.y, x.z);
    // A particle inside the grid (or exactly on its x == L or y == L face,
    // P17) has a key below the padding key.  One beyond is an error (as in
    // the reference, src/sorting.cuh:39-43); it is kept inside the
    // simulation's key range so that it cannot reach a neighbour's cells.
    if (lk >= p.pad_key - p.cell_base) {
        f.outside = 1;
        lk = p.pad_key - p.cell_base - 1;
    }
    const uint32_t key = p.cell_base + lk;
    if (x.x != x.x || x.y != x.y || x.z != x.z) f.nonfinite = 1;
    f.changed = key != key_prev;
    a.key[i] = key;
    if (p.vl_enabled) {
        // Candidate lists stay valid while every particle is within
        // skin/2 of where the lists were built.
        const double dx = x.x - xb.x, dy = x.y - xb.y, dz = x.z - xb.z;
        f.moved = !(dx * dx + dy * dy + dz * dz <= p.vl_trigger2);
    }
}

// The previous key and the list-build position of particle i, read ahead
// of the stores that precede hash_particle_from in the integration pass.
__device__ __forceinline__ void hash_inputs(CParams& p, const DevArrays& a, uint32_t i,
                                            uint32_t& key_prev, double4& xb) {
    key_prev = a.key_prev[i];
    xb = p.vl_enabled ? a.x_build[i] : make_double4(0.0, 0.0, 0.0, 0.0);
}

__device__ __forceinline__ void hash_particle(CParams& p, const DevArrays& a, uint32_t i,
                                              double Lz, const double4 x, HashFlags& f) {
    uint32_t key_prev;
    double4 xb;
    hash_inputs(p, a, i, key_prev, xb);
    hash_particle_from(p, a, i, Lz, x, key_prev, xb, f);
}

// Called by every work-item of the wavefront: at most one atomic per
// wavefront, and only when something happened.
__device__ __forceinline__ void publish_hash(DevLoop* lp, DevGlobal* g, uint32_t slot,
                                             const HashFlags& f) {
    const unsigned long long anych = __ballot(f.changed);
    const unsigned long long anymv = __ballot(f.moved);
    const unsigned long long anynf = __ballot(f.nonfinite);
    const unsigned long long anyout = __ballot(f.outside);
    if ((threadIdx.x & 63) == 0) {
        if (anyout) atomicOr(&lp->error_bits, (uint32_t)DFLAG_KEY_RANGE);
        if (anych) {
            atomicOr(&lp->changed[slot], 1u);
            atomicOr(&g->changed[slot], 1u);
        }
        if (anych | anymv) atomicOr(&lp->rebuild[slot], 1u);
        if (anynf) atomicOr(&lp->error_bits, (uint32_t)DFLAG_NONFINITE);
    }
}

// One workgroup per 256-particle chunk; every wavefront of a workgroup
// belongs to one simulation.
__global__ __launch_bounds__(BLOCK) void k_hash(const DevArrays a) {
    CParams& p = chunk_params(a, blockIdx.x);
    DevLoop* lp = a.loop + p.sim;
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    HashFlags f;
    if (i - p.base < p.np) hash_particle(p, a, i, grid_Lz(p, a), a.x[i], f);
    publish_hash(lp, a.glob, step_slot(lp), f);
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort with digits of `bits` bits (4, 6 or 8: at most
// 256 bins).  Tile t covers keys [t*TILE, (t+1)*TILE) in rounds of 256:
// element order inside a tile is (round, thread), which is the global
// index order, so ranking each round after all earlier rounds keeps the
// sort stable.  The stages are written per tile so that the same code runs
// as one launch per stage or inside the one-workgroup re-sort.

__device__ __forceinline__ void radix_hist_tile(const uint32_t* __restrict__ kin, uint32_t n,
                                                int shift, int bits,
                                                uint32_t* __restrict__ hist, uint32_t ntiles,
                                                uint32_t tile) {
    __shared__ uint32_t h[256];
    const uint32_t bins = 1u << bits, mask = bins - 1u;
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = tile * RADIX_TILE;
#pragma unroll
    for (int j = 0; j < RADIX_ITEMS; ++j) {
        const uint32_t e = base + j * BLOCK + threadIdx.x;
        if (e < n) atomicAdd(&h[(kin[e] >> shift) & mask], 1u);
    }
    __syncthreads();
    if (threadIdx.x < bins) hist[threadIdx.x * ntiles + tile] = h[threadIdx.x];
}

// Exclusive scan in place of the digit-major histogram (256 x ntiles) by
// one workgroup of T work-items.
template <int T>
__device__ __forceinline__ void radix_scan_block(uint32_t* __restrict__ hist, uint32_t total) {
    __shared__ uint32_t sh[T];
    const uint32_t t = threadIdx.x;
    const uint32_t chunk = (total + T - 1) / T;
    const uint32_t b0 = t * chunk;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < chunk; ++k) {
        const uint32_t e = b0 + k;
        if (e < total) sum += hist[e];
    }
    sh[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < (uint32_t)T; off <<= 1) {
        const uint32_t v = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    uint32_t run = sh[t] - sum;  // exclusive prefix of this chunk
    for (uint32_t k = 0; k < chunk; ++k) {
        const uint32_t e = b0 + k;
        if (e < total) {
            const uint32_t v = hist[e];
            hist[e] = run;
            run += v;
        }
    }
    __syncthreads();
}

// Without vin the values are vbase + the element index.
__device__ __forceinline__ void radix_scatter_tile(const uint32_t* __restrict__ kin,
                                                   const uint32_t* __restrict__ vin,
                                                   uint32_t* __restrict__ kout,
                                                   uint32_t* __restrict__ vout, uint32_t n,
                                                   int shift, int bits,
                                                   const uint32_t* __restrict__ hist,
                                                   uint32_t ntiles, uint32_t tile,
                                                   uint32_t vbase) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t whist[BLOCK / 64][256];
    const uint32_t bins = 1u << bits, mask = bins - 1u;
    const uint32_t t = threadIdx.x;
    const uint32_t w = t >> 6;
    const uint32_t lane = t & 63u;
    running[t] = t < bins ? hist[t * ntiles + tile] : 0u;
    const uint32_t base = tile * RADIX_TILE;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int j = 0; j < RADIX_ITEMS; ++j) {
#pragma unroll
        for (int q = 0; q < BLOCK / 64; ++q) whist[q][t] = 0u;
        __syncthreads();
        const uint32_t e = base + j * BLOCK + t;
        const bool valid = e < n;
        const uint32_t k = valid ? kin[e] : 0u;
        const uint32_t v = valid ? (vin ? vin[e] : vbase + e) : 0u;
        const uint32_t d = (k >> shift) & mask;
        unsigned long long peers = __ballot(valid);
        for (int bit = 0; bit < bits; ++bit) {
            const bool on = (d >> bit) & 1u;
            const unsigned long long b = __ballot(on);
            peers &= on ? b : ~b;
        }
        const uint32_t rank = __popcll(peers & lt);
        if (valid && rank == 0) whist[w][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = running[d] + rank;
            for (uint32_t q = 0; q < w; ++q) off += whist[q][d];
            kout[off] = k;
            vout[off] = v;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int q = 0; q < BLOCK / 64; ++q) add += whist[q][t];
        running[t] += add;
        __syncthreads();
    }
}

// Buffers of pass `pass` out of `passes` over the elements [b, ...): the
// last pass lands in (skey, perm); pass 0 reads the hash keys with the
// identity as values.
struct RadixPass {
    const uint32_t* kin;
    const uint32_t* vin;
    uint32_t* kout;
    uint32_t* vout;
};
__host__ __device__ inline RadixPass radix_pass(const DevArrays& a, int pass, int passes,
                                                uint32_t b) {
    RadixPass r;
    const bool to_final = ((passes - 1 - pass) % 2) == 0;
    if (pass == 0) {
        r.kin = a.key + b;
        r.vin = nullptr;
    } else {
        const bool prev_final = ((passes - pass) % 2) == 0;
        r.kin = (prev_final ? a.skey : a.key_tmp) + b;
        r.vin = (prev_final ? a.perm : a.val_tmp) + b;
    }
    r.kout = (to_final ? a.skey : a.key_tmp) + b;
    r.vout = (to_final ? a.perm : a.val_tmp) + b;
    return r;
}

// Every per-stage kernel returns at once when no particle changed cell this
// step: the stable order of (cell, index) is then unchanged.
__global__ __launch_bounds__(BLOCK) void k_radix_hist(const DevGlobal* __restrict__ g,
                                                      const DevLoop* __restrict__ lp,
                                                      const uint32_t* __restrict__ kin,
                                                      uint32_t n, int shift, int bits,
                                                      uint32_t* __restrict__ hist,
                                                      uint32_t ntiles) {
    if (!sort_needed(g, lp)) return;
    radix_hist_tile(kin, n, shift, bits, hist, ntiles, blockIdx.x);
}

__global__ __launch_bounds__(1024) void k_radix_scan(const DevGlobal* __restrict__ g,
                                                     const DevLoop* __restrict__ lp,
                                                     uint32_t* __restrict__ hist,
                                                     uint32_t total) {
    if (!sort_needed(g, lp)) return;
    radix_scan_block<1024>(hist, total);
}

__global__ __launch_bounds__(BLOCK) void k_radix_scatter(
    const DevGlobal* __restrict__ g, const DevLoop* __restrict__ lp,
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, int shift, int bits,
    const uint32_t* __restrict__ hist, uint32_t ntiles) {
    if (!sort_needed(g, lp)) return;
    radix_scatter_tile(kin, vin, kout, vout, n, shift, bits, hist, ntiles, blockIdx.x, 0u);
}

// ---------------------------------------------------------------------------
// Cells and reorder

__global__ __launch_bounds__(BLOCK) void k_fill_u32(const DevGlobal* __restrict__ g,
                                                    const DevLoop* __restrict__ lp,
                                                    uint32_t* __restrict__ dst, uint32_t n,
                                                    uint32_t v) {
    if (g && !sort_needed(g, lp)) return;
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) dst[i] = v;
}

// reorderArrays (src/sorting.cuh:56-132) for sorted index i of simulation
// p: cell ranges, the sorted copies, and the inverse permutation.  Keys are
// global (cell_base + the simulation's own key); only keys of real cells
// get a range (a key beyond the grid, or of a padding particle, does not).
__device__ __forceinline__ void cells_one(CParams& p, const DevArrays& a, uint32_t i,
                                          int sorted_bits, uint32_t par, bool copies) {
    const uint32_t key = a.skey[i];
    if (sorted_bits < 32 && (key >> sorted_bits) != 0u)
        atomicOr(&a.loop[p.sim].error_bits, (uint32_t)DFLAG_KEY_RANGE);
    const uint32_t lk = key - p.cell_base;
    if (i == 0 || key != a.skey[i - 1]) {
        if (lk < p.ncells) a.cell_start[key] = i;
        if (i > 0) {
            const uint32_t pk = a.skey[i - 1];
            if (pk - p.cell_base < p.ncells) a.cell_end[pk] = i;
        }
    }
    if (i == p.base + p.np - 1 && lk < p.ncells) a.cell_end[key] = i + 1;
    const uint32_t o = a.perm[i];
    if (copies) {
        a.xs[par][i] = a.x[o];
        a.vs[par][i] = a.vel[o];
        a.ws[par][i] = a.angvel[o];
    }
    a.rank[o] = i;
    a.key_prev[o] = key;
}

// Sorted ranges coincide with particle ranges (each simulation's keys,
// padding included, sit between its neighbours' keys), so chunk i / 256 of
// the sorted order belongs to the same simulation as chunk i / 256 of the
// particles.
__global__ __launch_bounds__(BLOCK) void k_cells(const DevArrays a, int sorted_bits) {
    if (!sort_needed(a.glob, a.loop)) return;
    CParams& p = chunk_params(a, blockIdx.x);
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i == p.base) {
        atomicAdd(&a.loop[p.sim].sorts, 1u);
        a.loop[p.sim].order_changed = 1u;
    }
    cells_one(p, a, i, sorted_bits, cur_set(a.loop + p.sim), true);
}

// ---------------------------------------------------------------------------
// Contact models (src/contactmodels.cuh)

struct Acc {
    d3 F, T;
    double es_dot, ev_dot, p;
};

// What one particle-particle contact adds to the accumulators, kept apart so
// that contacts evaluated on different lanes are added in slot order.  The
// energy rates get up to two separate additions per contact (the normal and
// the tangential dissipation), which stay two additions.
struct Contrib {
    d3 F, T;
    double p, ev0, ev1, es;
    int nev, nes;
};

__device__ __forceinline__ void contrib_init(Contrib& c) {
    c.nev = 0;
    c.nes = 0;
}
__device__ __forceinline__ void add_ev(Contrib& c, double v) {
    if (c.nev == 0) c.ev0 = v; else c.ev1 = v;
    ++c.nev;
}
__device__ __forceinline__ void add_es(Contrib& c, double v) {
    c.es = v;
    c.nes = 1;
}
__device__ __forceinline__ void apply(Acc& A, const Contrib& c) {
    A.F += c.F;
    A.T += c.T;
    A.p += c.p;
    if (c.nev > 0) A.ev_dot += c.ev0;
    if (c.nev > 1) A.ev_dot += c.ev1;
    if (c.nes > 0) A.es_dot += c.es;
}

// contactLinear_wall, src/contactmodels.cuh:10-117.
__device__ __forceinline__ double contact_linear_wall(CParams& p, Acc& A, double ra,
                                                   const double4 vt, const double4 at, d3 n,
                                                   double delta, double wvel) {
    const d3 vel_linear = xyz(vt);
    const d3 angvel = xyz(at);
    const d3 vel = vel_linear + (ra + delta / 2.0) * cross(n, angvel) + wvel;
    const double vel_n = dot(vel_linear, n);
    const d3 vel_t = vel - n * (dot(n, vel));
    const double vel_t_length = length(vel_t);
    double k_n = p.k_n;
    if (p.E > .001) k_n = 3.141592654 / 2.0 * p.E * ra;
    const d3 f_n = fmax(0.0, -k_n * delta - p.gamma_wn * vel_n) * n;
    const double f_n_length = length(f_n);
    A.ev_dot += p.gamma_wn * vel_n * vel_n;
    d3 f_t = mk3(0.0, 0.0, 0.0);
    if (vel_t_length > 0.0 && p.gamma_wt > 0.0) {
        const d3 f_t_visc = p.gamma_wt * vel_t;
        const double f_t_visc_length = length(f_t_visc);
        const double f_t_limit = p.mu_wd * f_n_length;
        if (f_t_visc_length < f_t_limit) {
            f_t = -1.0 * f_t_visc;
        } else {
            f_t = -f_t_limit * vel_t / vel_t_length;
            A.es_dot += length(length(f_t) * vel_t * p.dt) / p.dt;
        }
    }
    A.F += f_n + f_t;
    A.T += cross(-(ra + delta * 0.5) * n, f_t);
    A.p += f_n_length / (4.0 * PI * ra * ra);
    return dot(f_n, n);
}

// The factor -kappa sqrt32(r_a r_b) of the contact laws' capillary term
// f_c = -kappa sqrt32(r_a r_b) n.  Without cohesion (kappa == 0) and with
// r_a r_b > 0 finite in FP32 the factor is a zero of the sign of -kappa
// whatever the root is, so the root is skipped (`zero` set); then every
// component of f_n - f_c equals f_n's up to the sign of a zero, which the
// squares in length() drop: |f_n - f_c| = |f_n|.
__device__ __forceinline__ double capillary_factor(double kappa, double ra, double rb,
                                                   bool& zero) {
    const double rr = ra * rb;
    zero = kappa == 0.0 && rr > 0.0 && rr <= 3.4028234663852886e38;
    return -kappa * (zero ? 1.0 : sqrt32(rr));
}

// contactLinear (model 2), src/contactmodels.cuh:244-411.  Returns the new
// tangential displacement (stored by the caller).
__device__ __forceinline__ d3 contact_linear(CParams& p, Contrib& C, const double4 vel_a,
                                             d3 angvel_a, const double4 vel_b, d3 angvel_b,
                                             double ra, double rb, d3 x, double x_length,
                                             double delta, const double4 dt0) {
    const d3 delta_t0_uncor = mk3(dt0.x, dt0.y, dt0.z);
    const d3 n = x / x_length;
    const d3 vel_linear = mk3(vel_a.x - vel_b.x, vel_a.y - vel_b.y, vel_a.z - vel_b.z);
    const d3 vel = vel_linear + (ra + delta / 2.0) * cross(n, angvel_a) +
                   (rb + delta / 2.0) * cross(n, angvel_b);
    const double vel_n = -dot(vel_linear, n);
    const d3 vel_t = vel - n * dot(n, vel);
    const double vel_t_length = length(vel_t);
    const d3 delta_t0 = delta_t0_uncor - (n * dot(n, delta_t0_uncor));
    // length(delta_t0) > 0 exactly when the FP32 square is (the root of a
    // positive FP32 value is positive, of zero zero, of NaN NaN).
    const bool delta_t0_nonzero = (float)dot(delta_t0, delta_t0) > 0.0f;
    d3 delta_t = delta_t0;  // the value stored when neither branch assigns it
    double k_n = p.k_n, k_t = p.k_t;
    if (p.E > .001) {
        k_n = 3.141592654 / 2.0 * p.E * (ra + rb) / 2.;
        k_t = k_n;
    }
    const d3 f_n = fmax(0.0, -k_n * delta + p.gamma_n * vel_n) * n;
    const double f_n_length = length(f_n);
    add_ev(C, 0.5 * p.gamma_n * vel_n * vel_n);
    bool fc_zero;
    const d3 f_c = capillary_factor(p.kappa, ra, rb, fc_zero) * n;
    d3 f_t = mk3(0.0, 0.0, 0.0);
    if (delta_t0_nonzero || vel_t_length > 0.0) {
        delta_t = delta_t0 + vel_t * p.dt;
        const d3 f_t_elast = -k_t * delta_t;
        const d3 f_t_visc = -p.gamma_t * vel_t;
        f_t = f_t_elast + f_t_visc;
        const double f_t_length = length(f_t);
        const double fnc_length = fc_zero ? f_n_length : length(f_n - f_c);
        const double f_t_limit = p.mu_s * fnc_length;
        add_ev(C, 0.5 * p.gamma_t * vel_t_length * vel_t_length);
        if (f_t_length > f_t_limit) {
            const d3 t = f_t / f_t_length;
            f_t = f_t_limit * t;
            delta_t = -1.0 / k_t * (p.mu_d * fnc_length * t + p.gamma_t * vel_t);
            add_es(C, 0.5 * length(length(f_t) * vel_t * p.dt) / p.dt);
        }
    }
    C.F = f_n + f_t + f_c;
    C.T = cross(-(ra + delta * 0.5) * n, f_t);
    C.p = f_n_length / (4.0 * PI * ra * ra);
    return delta_t;
}

// contactHertz (model 3), src/contactmodels.cuh:417-600, with powers of
// |delta| through correctly rounded FP64 square roots and a repulsive
// elastic normal term (the reference's powf of a negative overlap is NaN).
__device__ __forceinline__ d3 contact_hertz(CParams& p, Contrib& C, const double4 vel_a,
                                            d3 angvel_a, const double4 vel_b, d3 angvel_b,
                                            double ra, double rb, d3 x_ab, double x_ab_length,
                                            double delta, const double4 dt0) {
    const d3 delta_t0_uncor = mk3(dt0.x, dt0.y, dt0.z);
    const d3 n_ab = x_ab / x_ab_length;
    const d3 vel_ab_linear = mk3(vel_a.x - vel_b.x, vel_a.y - vel_b.y, vel_a.z - vel_b.z);
    const d3 vel_ab = vel_ab_linear + (ra + delta / 2.0) * cross(n_ab, angvel_a) +
                      (rb + delta / 2.0) * cross(n_ab, angvel_b);
    const d3 angvel_ab = angvel_a - angvel_b;
    const double angvel_ab_length = length(angvel_ab);
    const double vel_n_ab = dot(vel_ab_linear, n_ab);
    const d3 vel_t_ab = vel_ab - (n_ab * dot(vel_ab, n_ab));
    const double vel_t_ab_length = length(vel_t_ab);
    const d3 delta_t0 = delta_t0_uncor - (n_ab * dot(delta_t0_uncor, n_ab));
    const bool delta_t0_nonzero = (float)dot(delta_t0, delta_t0) > 0.0f;  // length > 0
    d3 delta_t = delta_t0;
    double k_n = p.k_n, k_t = p.k_t;
    if (p.E > .001) {
        k_n = 3.141592654 / 2.0 * p.E * (ra + rb) / 2.;
        k_t = k_n;
    }
    const double ad = fabs(delta);
    const double p12 = sqrt_rn_f64(ad);
    const double p14 = sqrt_rn_f64(p12);
    const double p32 = ad * p12;
    d3 f_n = (k_n * p32 - p.gamma_n * p14 * vel_n_ab) * n_ab;
    add_ev(C, p.gamma_n * vel_n_ab * vel_n_ab);
    if (dot(f_n, n_ab) < 0.0) f_n = mk3(0.0, 0.0, 0.0);
    const double f_n_length = length(f_n);
    bool fc_zero;
    const d3 f_c = capillary_factor(p.kappa, ra, rb, fc_zero) * n_ab;
    d3 f_t = mk3(0.0, 0.0, 0.0);
    d3 T_res = mk3(0.0, 0.0, 0.0);
    if (delta_t0_nonzero || vel_t_ab_length > 0.0) {
        const d3 f_t_elast = -k_t * p12 * delta_t0;
        const d3 f_t_visc = -p.gamma_t * p14 * vel_t_ab;
        const double fnc_length = fc_zero ? f_n_length : length(f_n - f_c);
        double f_t_limit;
        if (vel_t_ab_length > (double)0.001f)
            f_t_limit = p.mu_d * fnc_length;
        else
            f_t_limit = p.mu_s * fnc_length;
        f_t = f_t_elast + f_t_visc;
        const double f_t_length = length(f_t);
        if (f_t_length > f_t_limit) {
            f_t *= f_t_limit / f_t_length;
            delta_t = (f_t + p.gamma_t * p14 * vel_t_ab) / (-k_t * p12);
            add_es(C, length(delta_t0 - delta_t) * k_t / p.dt);
        } else {
            delta_t = delta_t0 + vel_t_ab * p.dt;
        }
    }
    if (angvel_ab_length > 0.0) {
        T_res = -1.0 * fmin(p.gamma_r * ra * angvel_ab_length, p.mu_r * ra * f_n_length) *
                angvel_ab / angvel_ab_length;
    }
    C.F = f_n + f_t + f_c;
    C.T = -(ra + delta / 2.0) * cross(n_ab, f_t) + T_res;
    C.p = f_n_length / (4.0 * PI * ra * ra);
    return delta_t;
}

// contactLinearViscous (model 1), src/contactmodels.cuh:122-240.
__device__ __forceinline__ void contact_linear_viscous(CParams& p, Acc& A,
                                                       const double4 vel_a, d3 angvel_a,
                                                       const double4 vel_b, d3 angvel_b,
                                                       double ra, double rb, d3 x_ab,
                                                       double x_ab_length, double delta,
                                                       double kappa) {
    const d3 n_ab = x_ab / x_ab_length;
    const d3 vel_ab_linear = mk3(vel_a.x - vel_b.x, vel_a.y - vel_b.y, vel_a.z - vel_b.z);
    const d3 vel_ab = vel_ab_linear + (ra + delta / 2.0) * cross(n_ab, angvel_a) +
                      (rb + delta / 2.0) * cross(n_ab, angvel_b);
    const double vel_n_ab = dot(vel_ab_linear, n_ab);
    const d3 vel_t_ab = vel_ab - (n_ab * dot(vel_ab, n_ab));
    const double vel_t_ab_length = length(vel_t_ab);
    double k_n = p.k_n;
    if (p.E > .001) k_n = 3.141592654 / 2.0 * p.E * (ra + rb) / 2.;
    d3 f_n = fmax(0.0, -k_n * delta - p.gamma_n * vel_n_ab) * n_ab;
    if (dot(f_n, n_ab) < 0.0) f_n = mk3(0.0, 0.0, 0.0);
    const double f_n_length = length(f_n);
    bool fc_zero;
    const d3 f_c = capillary_factor(kappa, ra, rb, fc_zero) * n_ab;
    d3 f_t = mk3(0.0, 0.0, 0.0);
    if (vel_t_ab_length > 0.0) {
        const double f_t_visc = p.gamma_t * vel_t_ab_length;
        const double fnc_length = fc_zero ? f_n_length : length(f_n - f_c);
        double f_t_limit;
        if (vel_t_ab_length > (double)0.001f)
            f_t_limit = p.mu_d * fnc_length;
        else
            f_t_limit = p.mu_s * fnc_length;
        if (f_t_visc < f_t_limit)
            f_t = -f_t_visc * vel_t_ab / vel_t_ab_length;
        else
            f_t = -f_t_limit * vel_t_ab / vel_t_ab_length;
    }
    A.F += f_n + f_t + f_c;
    A.T += -(ra + delta / 2.0) * cross(n_ab, f_t);
    A.p += f_n_length / (4.0 * PI * ra * ra);
}

// capillaryCohesion_exp, src/cohesion.cuh:237-277 (FP32 sqrtf and expf as
// there; expf's last bit is not pinned across libraries).
__device__ __forceinline__ void capillary_cohesion_exp(CParams& p, Acc& A, double ra,
                                                       double rb, double delta, d3 x_ab,
                                                       double x_ab_length, double kappa) {
    const d3 n_ab = x_ab / x_ab_length;
    double r;
    if ((ra / rb) > (rb / ra)) r = ra / rb; else r = rb / ra;
    const double h = -sqrt32(r);
    const double R_har = (2.0 * ra * rb) / (ra + rb);
    const double R_geo = sqrt32(ra * rb);
    const double lambda = 0.9 * h * sqrt32(p.V_b / R_har);
    const d3 f_c = -kappa * R_geo * (double)expf((float)(-delta / lambda)) * n_ab;
    A.F += f_c;
}

// findDistMod, src/contactsearch.cuh:12-83.  Returns the image code (see
// CW_SHIFT_*) or -1 when the target cell lies outside the grid.
__device__ __forceinline__ int find_dist_mod(CParams& p, int& tx, int& ty, int tz) {
    const int nx = (int)p.num[0], ny = (int)p.num[1], nz = (int)p.num[2];
    int code = 0;
    if (p.periodic == 1) {
        if (tx < 0) { tx += nx; code |= 1; }
        if (tx >= nx) { tx -= nx; code |= 2; }
        if (ty < 0) { ty += ny; code |= 4; }
        if (ty >= ny) { ty -= ny; code |= 8; }
    }
    if (p.periodic == 2) {
        if (tx < 0) { tx += nx; code |= 1; }
        if (tx >= nx) { tx -= nx; code |= 2; }
        if (ty < 0 || ty >= ny) return -1;
    }
    if (p.periodic > 2 || p.periodic < 1) {
        if (tx < 0 || tx >= nx) return -1;
        if (ty < 0 || ty >= ny) return -1;
    }
    if (tz < 0 || tz >= nz) return -1;
    // A particle far outside the grid can leave tx/ty out of range after one
    // periodic shift; the reference then reads outside cellStart.  Such a
    // cell is skipped here.
    if (tx < 0 || tx >= nx || ty < 0 || ty >= ny) return -1;
    return code;
}

// The image shift of a code, exactly as findDistMod builds it from zero.
__device__ __forceinline__ d3 shift_of(CParams& p, uint32_t code) {
    d3 s = mk3(0.0, 0.0, 0.0);
    if (code & 1u) s.x = 0.0 + p.L[0];
    if (code & 2u) s.x = 0.0 - p.L[0];
    if (code & 4u) s.y = 0.0 + p.L[1];
    if (code & 8u) s.y = 0.0 - p.L[1];
    return s;
}

__device__ __forceinline__ uint32_t pack_word(uint32_t idx, uint32_t code) {
    // code bits: 1 (+Lx), 2 (-Lx), 4 (+Ly), 8 (-Ly)
    uint32_t w = idx;
    if (code & 1u) w |= 1u << CW_SHIFT_X;
    if (code & 2u) w |= 2u << CW_SHIFT_X;
    if (code & 4u) w |= 1u << CW_SHIFT_Y;
    if (code & 8u) w |= 2u << CW_SHIFT_Y;
    return w;
}
__device__ __forceinline__ uint32_t word_code(uint32_t w) {
    const uint32_t cx = (w >> CW_SHIFT_X) & 3u, cy = (w >> CW_SHIFT_Y) & 3u;
    return (cx == 1u ? 1u : 0u) | (cx == 2u ? 2u : 0u) | (cy == 1u ? 4u : 0u) |
           (cy == 2u ? 8u : 0u);
}

// A row word written in the order of perm_rows, mapped to the current
// sorted order through the partner's original index (after a re-sort).
__device__ __forceinline__ uint32_t word_to_current(const DevArrays& a, uint32_t w) {
    const uint32_t idx = w & CW_IDX_MASK;
    if (idx == EMPTY) return w;
    return (w & ~CW_IDX_MASK) | a.rank[a.perm_rows[idx]];
}

// A row word of the current sorted order mapped to the partner's original
// index (the adjoint's tape keeps original indices).
__device__ __forceinline__ uint32_t word_to_orig(const DevArrays& a, uint32_t w) {
    const uint32_t idx = w & CW_IDX_MASK;
    if (idx == EMPTY) return w;
    return (w & ~CW_IDX_MASK) | a.perm[idx];
}

// The state the contact passes read by sorted index: the step's sorted
// copies, or on the step after a one-workgroup re-sort (which leaves them
// in the previous order, copies_stale) the original-index arrays through
// perm.  The flag is uniform over a simulation.
struct SortedReader {
    const double4* __restrict__ x;
    const double4* __restrict__ v;
    const double4* __restrict__ w;
    const uint32_t* __restrict__ perm;
    bool stale;
    __device__ __forceinline__ SortedReader(const DevArrays& a, const DevLoop* lp) {
        stale = lp->copies_stale != 0u;
        const uint32_t cs = cur_set(lp);
        x = stale ? a.x : a.xs[cs];
        v = stale ? a.vel : a.vs[cs];
        w = stale ? a.angvel : a.ws[cs];
        perm = a.perm;
    }
    __device__ __forceinline__ uint32_t at(uint32_t i) const { return stale ? perm[i] : i; }
};

// ---------------------------------------------------------------------------
// topology + interact, one work-item per particle in sorted order
// (src/contactsearch.cuh:329-371 then :385-726).  The particle's contact
// row lives in LDS for the duration (33-word stride: no bank conflicts).

constexpr int ROW_STRIDE = NC + 1;
#ifndef CONTACT_MIN_BLOCKS
#define CONTACT_MIN_BLOCKS 4
#endif

// The contact-slot policy of findContactsInCell for one overlapping
// partner (src/contactsearch.cuh:265-299): the last empty slot, else the
// partner's own slot, else slot nc.  Returns true when nc exceeds NC, which
// ends the scan of the current cell.
__device__ __forceinline__ bool register_contact(uint32_t* row, uint32_t ib, uint32_t code,
                                                 int& nc, DevLoop* loop) {
    int cpos = nc;
    for (int i = 0; i < NC; ++i) {
        const uint32_t cidx = row[i] & CW_IDX_MASK;
        if (cidx == EMPTY) {
            cpos = i;
        } else if (cidx == ib) {
            cpos = i;
            break;
        }
    }
    // cpos == NC would write into the next particle's row
    // (src/contactsearch.cuh:286); that write is suppressed.
    if (cpos < NC) row[cpos] = pack_word(ib, code);
    ++nc;
    if (nc > NC) {
        atomicOr(&loop->error_bits, (uint32_t)DFLAG_CONTACT_OVERFLOW);
        return true;
    }
    return false;
}

// Walls (src/contactsearch.cuh:602-701): top, bottom, then the side walls
// (all four when periodic == 0, the y walls when periodic == 2), each
// spelled out so that nothing is indexed dynamically; then the particle's
// outputs at its original index o when `write` is set.
__device__ __forceinline__ void walls_and_outputs(CParams& p, const DevArrays& a, Acc& A,
                                                  const double4 xa, const double4 vel_a,
                                                  const double4 av_a, double Lz, uint32_t o,
                                                  int part) {
    const double ra = xa.w;
    double wf0 = 0.0, wf1 = 0.0, wf2 = 0.0, wf3 = 0.0, wf4 = 0.0;
    {
        double4 wn = make_double4(0.0, 0.0, -1.0, Lz);
        double wv = 0.0;
        if (p.nw >= 1) { wn = wall_nx(a, p.wall_base); wv = wall_mvfd(a, p.wall_base).y; }
        const double dw = wn.w - (xa.z + ra);
        if (dw < 0.0) wf0 = contact_linear_wall(p, A, ra, vel_a, av_a, xyz(wn), dw, wv);
    }
    {
        const double dw = xa.z - ra - p.origo[2];
        if (dw < 0.0)
            (void)contact_linear_wall(p, A, ra, vel_a, av_a, mk3(0.0, 0.0, 1.0), dw, 0.0);
    }
    if (p.periodic == 0) {
        {
            double4 wn = make_double4(-1.0, 0.0, 0.0, p.L[0]);
            double wv = 0.0;
            if (p.nw >= 2) { wn = wall_nx(a, p.wall_base + 1); wv = wall_mvfd(a, p.wall_base + 1).y; }
            const double dw = wn.w - (xa.x + ra);
            if (dw < 0.0) wf1 = contact_linear_wall(p, A, ra, vel_a, av_a, xyz(wn), dw, wv);
        }
        {
            double4 wn = make_double4(1.0, 0.0, 0.0, 0.0);
            double wv = 0.0;
            if (p.nw >= 3) { wn = wall_nx(a, p.wall_base + 2); wv = wall_mvfd(a, p.wall_base + 2).y; }
            const double dw = xa.x - ra - wn.w;
            if (dw < 0.0) wf2 = contact_linear_wall(p, A, ra, vel_a, av_a, xyz(wn), dw, wv);
        }
    }
    if (p.periodic == 0 || p.periodic == 2) {
        {
            double4 wn = make_double4(0.0, -1.0, 0.0, p.L[1]);
            double wv = 0.0;
            if (p.nw >= 4) { wn = wall_nx(a, p.wall_base + 3); wv = wall_mvfd(a, p.wall_base + 3).y; }
            const double dw = wn.w - (xa.y + ra);
            if (dw < 0.0) wf3 = contact_linear_wall(p, A, ra, vel_a, av_a, xyz(wn), dw, wv);
        }
        {
            double4 wn = make_double4(0.0, 1.0, 0.0, 0.0);
            double wv = 0.0;
            if (p.nw >= 5) { wn = wall_nx(a, p.wall_base + 4); wv = wall_mvfd(a, p.wall_base + 4).y; }
            const double dw = xa.y - ra - wn.w;
            if (dw < 0.0) wf4 = contact_linear_wall(p, A, ra, vel_a, av_a, xyz(wn), dw, wv);
        }
    }
    // Outputs at the original index, src/contactsearch.cuh:706-724: part
    // `part` of the particle's output record (all four with part == 4; -1
    // writes nothing).  The integration pass advances es and ev by them.
    double4* const rec = a.out + (size_t)o * 4;
    if (part == 0 || part == 4) rec[OUT_F] = make_double4(A.F.x, A.F.y, A.F.z, 0.0);
    if (part == 1 || part == 4) rec[OUT_T] = make_double4(A.T.x, A.T.y, A.T.z, 0.0);
    if (part == 2 || part == 4) rec[OUT_E] = make_double4(A.es_dot, A.ev_dot, A.p, wf0);
    if (part == 3 || part == 4) rec[OUT_W] = make_double4(wf1, wf2, wf3, wf4);
}

// es += es_dot dt and ev += ev_dot dt of the contact pass's outputs
// (src/contactsearch.cuh:716-717), for particle i.
__device__ __forceinline__ void advance_energies(CParams& p, const DevArrays& a,
                                                 uint32_t i) {
    const double4 e = a.out[(size_t)i * 4 + OUT_E];
    a.es[i] = a.es[i] + e.x * p.dt;
    a.ev[i] = a.ev[i] + e.y * p.dt;
}

// A particle's force on wall w, from its output record (walls beyond the
// top and the four side walls take none).
__device__ __forceinline__ double wall_force_pp(const DevArrays& a, uint32_t i, uint32_t w) {
    const double4* rec = a.out + (size_t)i * 4;
    if (w == 0) return rec[OUT_E].w;
    if (w >= OUT_NWALL) return 0.0;
    const double4 v = rec[OUT_W];
    return w == 1 ? v.x : w == 2 ? v.y : w == 3 ? v.z : v.w;
}

// The adjoint's record of the step's starting state: each particle's and
// the order, at its sorted index, and each simulation's walls by the
// work-item of its first particle.
__device__ __forceinline__ void tape_start(const DevArrays& a, CParams& p, uint32_t ia,
                                           uint32_t ia_orig, double4 x, double4 v, double4 w,
                                           bool writer) {
    if (!a.tape_x || !writer) return;
    a.tape_x[ia] = x;
    a.tape_v[ia] = v;
    a.tape_w[ia] = w;
    a.tape_perm[ia] = ia_orig;
    if (ia_orig == p.base)
        for (uint32_t k = p.wall_base; k < p.wall_base + p.nw; ++k) {
            a.tape_wnx[k] = a.w_nx[k];
            a.tape_wmv[k] = a.w_mvfd[k];
        }
}

// CM is the contact model (1, 2 or 3); one instantiation per model keeps
// only that model's registers live.
template <int CM>
__global__ __launch_bounds__(BLOCK, CONTACT_MIN_BLOCKS) void k_contact(const DevArrays a) {
    __shared__ uint32_t srow[BLOCK * ROW_STRIDE];
    const uint32_t blk = xcd_block(a, blockIdx.x, gridDim.x);
    CParams& p = chunk_params(a, blk);
    const uint32_t ia = blk * BLOCK + threadIdx.x;
    if (ia - p.base >= p.np) return;
    uint32_t* row = srow + threadIdx.x * ROW_STRIDE;
    DevLoop* lp = a.loop + p.sim;
    const uint32_t ia_orig = a.perm[ia];
    const SortedReader S(a, lp);
    const double4 xa = S.x[S.at(ia)];
    const double ra = xa.w;
    const double Lz = grid_Lz(p, a);
    const double4 vel_a = S.v[S.at(ia)];
    const double4 av_a = S.w[S.at(ia)];
    const d3 angvel_a = xyz(av_a);
    tape_start(a, p, ia, ia_orig, xa, vel_a, av_a, true);
    Acc A;
    A.F = mk3(0.0, 0.0, 0.0);
    A.T = mk3(0.0, 0.0, 0.0);
    A.es_dot = 0.0;
    A.ev_dot = 0.0;
    A.p = 0.0;
    // src/contactsearch.cuh:352-354
    const int gx = (int)floor((xa.x - p.origo[0]) / (p.L[0] / p.num[0]));
    const int gy = (int)floor((xa.y - p.origo[1]) / (p.L[1] / p.num[1]));
    const int gz = (int)floor((xa.z - p.origo[2]) / (Lz / p.num[2]));
    if constexpr (CM == 2 || CM == 3) {
        const uint4* grow = reinterpret_cast<const uint4*>(a.crow + (size_t)ia_orig * NC);
#pragma unroll
        for (int q = 0; q < NC / 4; ++q) {
            const uint4 v = grow[q];
            row[4 * q] = v.x; row[4 * q + 1] = v.y; row[4 * q + 2] = v.z; row[4 * q + 3] = v.w;
        }
        // After a re-sort the words are mapped to the new order.
        if (lp->order_changed != 0u)
            for (int i = 0; i < NC; ++i) row[i] = word_to_current(a, row[i]);
        // topology (findContactsInCell over the 27 cells, z-major, each cell
        // in ascending sorted index).  Three ways to visit the same candidate
        // sequence: the full scan that also rebuilds this particle's
        // candidate list, the full scan alone (list overflowed), or the list.
        int nc = 0;
        const bool rebuild =
            !p.vl_enabled || lp->rebuild[step_slot(lp)] != 0u || lp->force_rebuild != 0u;
        const uint2 vmeta = rebuild ? make_uint2(0u, 0u) : a.vcount[ia];
        const uint32_t vcnt = vmeta.x == VL_OVERFLOW ? VL_OVERFLOW : (vmeta.x & VC_COUNT_MASK);
        // This kernel keeps no hit mask: a valid one would go stale.
        if (!rebuild && vcnt != VL_OVERFLOW && (vmeta.x & VC_HP_VALID) != 0u)
            a.vcount[ia] = make_uint2(vcnt, 0u);
        if (rebuild || vcnt == VL_OVERFLOW) {
            const bool build = p.vl_enabled && rebuild;
            uint32_t nl = 0;
            uint2* vrow = a.vlist + (size_t)ia * NV;
            for (int dz = -1; dz < 2; ++dz)
                for (int dy = -1; dy < 2; ++dy)
                    for (int dx = -1; dx < 2; ++dx) {
                        int tx = gx + dx, ty = gy + dy;
                        const int tz = gz + dz;
                        const int code = find_dist_mod(p, tx, ty, tz);
                        if (code < 0) continue;
                        const uint32_t cell = p.cell_base + (uint32_t)tx +
                                              (uint32_t)ty * p.num[0] +
                                              (p.num[0] * p.num[1]) * (uint32_t)tz;
                        const uint32_t start = a.cell_start[cell];
                        if (start == 0xffffffffu) continue;
                        const uint32_t end = a.cell_end[cell];
                        const d3 dm = shift_of(p, (uint32_t)code);
                        const uint32_t celloff = (uint32_t)((dz + 1) * 9 + (dy + 1) * 3 + (dx + 1));
                        bool cell_done = false;
                        for (uint32_t ib = start; ib < end; ++ib) {
                            if (ib == ia) continue;
                            const double4 xb = S.x[S.at(ib)];
                            const double rb = xb.w;
                            d3 xab = mk3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z);
                            xab += dm;
                            const double dd = dot(xab, xab);
                            const double Sab = ra + rb;
                            if (build) {
                                const double Sv = Sab + p.vl_skin;
                                if (dd < Sv * Sv) {
                                    if (nl < (uint32_t)NV)
                                        vrow[nl] = make_uint2(ib | (celloff << VL_CELL_SHIFT),
                                                              pack_word(ib, (uint32_t)code));
                                    ++nl;
                                }
                            }
                            if (cell_done) continue;
                            // Conservative reject: len32 > S whenever
                            // dd >= S^2 (1 + 2^-16) (the two FP32 roundings
                            // move the root by < 2^-23 relative).
                            if (dd >= Sab * Sab * 1.0000152587890625) continue;
                            const double len = sqrt32(dd);
                            if (len - Sab < 0.0)
                                cell_done = register_contact(row, ib, (uint32_t)code, nc, lp);
                        }
                    }
            if (build) {
                a.vcount[ia] = make_uint2(nl <= p.vl_cap ? nl : VL_OVERFLOW, 0u);
                a.x_build[ia_orig] = xa;
            }
        } else {
            const uint2* vrow = a.vlist + (size_t)ia * NV;
            uint32_t last_cell = 0xffffffffu;
            bool cell_done = false;
            for (uint32_t k = 0; k < vcnt; ++k) {
                const uint2 e = vrow[k];
                const uint32_t celloff = e.x >> VL_CELL_SHIFT;
                if (celloff != last_cell) {
                    last_cell = celloff;
                    cell_done = false;
                }
                if (cell_done) continue;
                const double4 xb = S.x[S.at(e.x & VL_IB_MASK)];
                const double rb = xb.w;
                const uint32_t code = word_code(e.y);
                d3 xab = mk3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z);
                xab += shift_of(p, code);
                const double dd = dot(xab, xab);
                const double Sab = ra + rb;
                if (dd >= Sab * Sab * 1.0000152587890625) continue;
                const double len = sqrt32(dd);
                if (len - Sab < 0.0)
                    cell_done = register_contact(row, e.y & CW_IDX_MASK, code, nc, lp);
            }
        }
        // interact: the contact loop, src/contactsearch.cuh:484-559.
        const bool zero_empty = lp->zero_empty != 0u;
        for (int i = 0; i < NC; ++i) {
            const size_t mp = (size_t)ia_orig * NC + i;
            const uint32_t wrd = row[i];
            const uint32_t rb_s = wrd & CW_IDX_MASK;
            if (rb_s != EMPTY) {
                const d3 dm = shift_of(p, word_code(wrd));
                const uint32_t rb_j = S.at(rb_s);
                const double4 xb = S.x[rb_j];
                const double rb = xb.w;
                const d3 xab = mk3(xa.x - xb.x + dm.x, xa.y - xb.y + dm.y, xa.z - xb.z + dm.z);
                const double len = length(xab);
                const double dn = len - (ra + rb);
                if (dn < 0.0) {
                    const double4 vel_b = S.v[rb_j];
                    const d3 angvel_b = xyz(S.w[rb_j]);
                    const double4 dt0 = a.delta_t[mp];
                    if (a.tape_dt0) {
                        a.tape_dt0[mp] = dt0;
                        a.tape_srt[mp] = rb_s;
                    }
                    d3 nd;
                    Contrib C;
                    contrib_init(C);
                    if constexpr (CM == 2)
                        nd = contact_linear(p, C, vel_a, angvel_a, vel_b, angvel_b, ra, rb, xab,
                                            len, dn, dt0);
                    else
                        nd = contact_hertz(p, C, vel_a, angvel_a, vel_b, angvel_b, ra, rb, xab,
                                           len, dn, dt0);
                    apply(A, C);
                    a.delta_t[mp] = make_double4(nd.x, nd.y, nd.z, 0.0);
                } else {
                    row[i] = EMPTY;
                    a.delta_t[mp] = make_double4(0.0, 0.0, 0.0, 0.0);
                }
            } else if (zero_empty) {
                // Empty slots hold a zero displacement once every slot has been
                // through one interact pass (src/contactsearch.cuh:554-558);
                // only the first pass after the history was loaded writes them.
                a.delta_t[mp] = make_double4(0.0, 0.0, 0.0, 0.0);
            }
        }
        uint4* wrow = reinterpret_cast<uint4*>(a.crow + (size_t)ia_orig * NC);
#pragma unroll
        for (int q = 0; q < NC / 4; ++q)
            wrow[q] = make_uint4(row[4 * q], row[4 * q + 1], row[4 * q + 2], row[4 * q + 3]);
        if (a.tape_rows)
            for (int i = 0; i < NC; ++i)
                a.tape_rows[(size_t)ia_orig * NC + i] = word_to_orig(a, row[i]);
    } else if constexpr (CM == 1) {
        // findAndProcessContactsInCell, src/contactsearch.cuh:91-186.
        for (int dz = -1; dz < 2; ++dz)
            for (int dy = -1; dy < 2; ++dy)
                for (int dx = -1; dx < 2; ++dx) {
                    int tx = gx + dx, ty = gy + dy;
                    const int tz = gz + dz;
                    const int code = find_dist_mod(p, tx, ty, tz);
                    if (code < 0) continue;
                    const uint32_t cell = p.cell_base + (uint32_t)tx + (uint32_t)ty * p.num[0] +
                                          (p.num[0] * p.num[1]) * (uint32_t)tz;
                    const uint32_t start = a.cell_start[cell];
                    if (start == 0xffffffffu) continue;
                    const uint32_t end = a.cell_end[cell];
                    const d3 dm = shift_of(p, (uint32_t)code);
                    for (uint32_t ib = start; ib < end; ++ib) {
                        if (ib == ia) continue;
                        const uint32_t jb = S.at(ib);
                        const double4 xb = S.x[jb];
                        const double rb = xb.w;
                        const double kappa = p.kappa;
                        d3 xab = mk3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z);
                        xab += dm;
                        const double len = length(xab);
                        const double delta = len - (ra + rb);
                        if (delta < 0.0) {
                            contact_linear_viscous(p, A, vel_a, angvel_a, S.v[jb],
                                                   xyz(S.w[jb]), ra, rb, xab, len,
                                                   delta, kappa);
                        } else if (delta < p.db) {
                            capillary_cohesion_exp(p, A, ra, rb, delta, xab, len, kappa);
                        }
                    }
                }
    }
    walls_and_outputs(p, a, A, xa, vel_a, av_a, Lz, ia_orig, 4);
}

// ---------------------------------------------------------------------------
// integrate (TY3), src/integration.cuh:14-305, one work-item per particle in
// original order (the reference's sorted-order pass reads copies of the
// same values; every output is per particle).

// One particle's step from its force, torque and state (x, vel, angvel
// carry radius and fixvel in w); writes the new state at its original index
// and the next step's sorted copies at sorted index `r`.
__device__ __forceinline__ double4 integrate_particle(CParams& p, const DevArrays& a,
                                                      uint32_t orig, const double4 force,
                                                      const double4 torque, const double4 x,
                                                      const double4 vel, const double4 angvel,
                                                      uint32_t r) {
    const DevLoop* lp = a.loop + p.sim;
    const double t_now = lp->t;
    const int64_t iter = lp->iter;
    // Rate-state switch decided by the host loop before integrate
    // (src/device.cu:2454-2462).
    int cvs = 0;
    if (t_now >= 10.0 && t_now < 15.0 && lp->velocity_state == 1) cvs = 1;
    else if (t_now >= 15.0 && lp->velocity_state == 2) cvs = -1;
    const double Lz = grid_Lz(p, a);
    const double4 angpos = a.angpos[orig];
    double4 xyzsum = a.xyzsum[orig];
    const int wall0mode = p.nw > 0 ? p.wmode[0] : 0;
    const double wall0mass = p.nw > 0 ? wall_mvfd(a, p.wall_base).x : 0.0;
    double tau_eff_x = 0.0;
    if (p.nw > 0 && wall0mode == 3) tau_eff_x = a.tau_eff[p.sim];
    double4 acc0, angacc0;
    if (iter == 0) {
        acc0 = make_double4(0.0, 0.0, 0.0, 0.0);
        angacc0 = make_double4(0.0, 0.0, 0.0, 0.0);
    } else {
        acc0 = a.acc[orig];
        angacc0 = a.angacc[orig];
    }
    const double radius = x.w;
    const double dt = p.dt;
    const double m = 4.0 / 3.0 * PI * radius * radius * radius * p.rho;
    double4 acc = make_double4(0.0, 0.0, 0.0, 0.0);
    acc.x = force.x / m + p.g[0];
    acc.y = force.y / m + p.g[1];
    acc.z = force.z / m + p.g[2];
    double4 angacc = make_double4(0.0, 0.0, 0.0, 0.0);
    angacc.x = torque.x * 1.0 / (2.0 / 5.0 * m * radius * radius);
    angacc.y = torque.y * 1.0 / (2.0 / 5.0 * m * radius * radius);
    angacc.z = torque.z * 1.0 / (2.0 / 5.0 * m * radius * radius);
    if (wall0mode == 3 && vel.w > 0.0001 && x.z > Lz * 0.5) {
        const double f_tau_x = (p.tau_x + tau_eff_x) * p.L[0] * p.L[1];
        acc.x = f_tau_x / wall0mass;
        acc.y = 0.0;
        acc.z -= p.g[2];
        angacc = make_double4(0.0, 0.0, 0.0, 0.0);
    } else if (vel.w > 0.0001 && vel.w < 10.0) {
        acc.x = 0.0;
        acc.y = 0.0;
        acc.z -= p.g[2];
        angacc = make_double4(0.0, 0.0, 0.0, 0.0);
    }
    if (vel.w >= 10.0) {
        acc.y = 0.0;
        angacc.x = 0.0;
        angacc.z = 0.0;
    }
    if (vel.w < -0.0001) {
        acc.x = 0.0;
        acc.y = 0.0;
        acc.z = 0.0;
        angacc = make_double4(0.0, 0.0, 0.0, 0.0);
    }
    const d3 dacc_dt = mk3((acc.x - acc0.x) / dt, (acc.y - acc0.y) / dt, (acc.z - acc0.z) / dt);
    const d3 dangacc_dt = mk3((angacc.x - angacc0.x) / dt, (angacc.y - angacc0.y) / dt,
                              (angacc.z - angacc0.z) / dt);
    double4 x_new, vel_new, angpos_new, angvel_new;
    x_new.x = x.x + vel.x * dt + 0.5 * acc.x * dt * dt + 1.0 / 6.0 * dacc_dt.x * dt * dt * dt;
    x_new.y = x.y + vel.y * dt + 0.5 * acc.y * dt * dt + 1.0 / 6.0 * dacc_dt.y * dt * dt * dt;
    x_new.z = x.z + vel.z * dt + 0.5 * acc.z * dt * dt + 1.0 / 6.0 * dacc_dt.z * dt * dt * dt;
    x_new.w = x.w;
    vel_new.x = vel.x + acc.x * dt + 0.5 * dacc_dt.x * dt * dt;
    vel_new.y = vel.y + acc.y * dt + 0.5 * dacc_dt.y * dt * dt;
    vel_new.z = vel.z + acc.z * dt + 0.5 * dacc_dt.z * dt * dt;
    vel_new.w = vel.w;
    angpos_new.x = angpos.x + angvel.x * dt + 0.5 * angacc.x * dt * dt +
                   1.0 / 6.0 * dangacc_dt.x * dt * dt * dt;
    angpos_new.y = angpos.y + angvel.y * dt + 0.5 * angacc.y * dt * dt +
                   1.0 / 6.0 * dangacc_dt.y * dt * dt * dt;
    angpos_new.z = angpos.z + angvel.z * dt + 0.5 * angacc.z * dt * dt +
                   1.0 / 6.0 * dangacc_dt.z * dt * dt * dt;
    angpos_new.w = angpos.w;
    angvel_new.x = angvel.x + angacc.x * dt + 0.5 * dangacc_dt.x * dt * dt;
    angvel_new.y = angvel.y + angacc.y * dt + 0.5 * dangacc_dt.y * dt * dt;
    angvel_new.z = angvel.z + angacc.z * dt + 0.5 * dangacc_dt.z * dt * dt;
    angvel_new.w = angvel.w;
    xyzsum.x += vel.x * dt + 0.5 * acc.x * dt * dt + 1.0 / 6.0 * dacc_dt.x * dt * dt * dt;
    xyzsum.y += vel.y * dt + 0.5 * acc.y * dt * dt + 1.0 / 6.0 * dacc_dt.y * dt * dt * dt;
    xyzsum.z += vel.z * dt + 0.5 * acc.z * dt * dt + 1.0 / 6.0 * dacc_dt.z * dt * dt * dt;
    // Periodic wrap, src/integration.cuh:267-283 (note `> L`, P17).
    if (p.periodic == 1) {
        if (x_new.x < p.origo[0]) x_new.x += p.L[0];
        if (x_new.x > p.L[0]) x_new.x -= p.L[0];
        if (x_new.y < p.origo[1]) x_new.y += p.L[1];
        if (x_new.y > p.L[1]) x_new.y -= p.L[1];
    } else if (p.periodic == 2) {
        if (x_new.x < p.origo[0]) x_new.x += p.L[0];
        if (x_new.x > p.L[0]) x_new.x -= p.L[0];
    }
    if (vel.w > 5.0 && vel.w < 10.0) {
        if (cvs == 1) vel_new.x *= 10.0;
        else if (cvs == -1) vel_new.x /= 10.0;
    }
    a.xyzsum[orig] = xyzsum;
    a.acc[orig] = acc;
    a.angacc[orig] = angacc;
    a.angvel[orig] = angvel_new;
    a.vel[orig] = vel_new;
    a.angpos[orig] = angpos_new;
    a.x[orig] = x_new;
    // Keep the sorted copies current: valid for the next step whenever the
    // order does not change (otherwise the next sort rebuilds them).  The
    // set follows the loop counter read above (cur_set).
    const uint32_t nxt = (uint32_t)(iter & 1) ^ 1u;
    a.xs[nxt][r] = x_new;
    a.vs[nxt][r] = vel_new;
    a.ws[nxt][r] = angvel_new;
    return x_new;
}

// integrate_particle for the four lanes of a contact group, one vector
// component per lane (lane q: x, y, z; lane 3 carries the w components of
// the next step's sorted copies): the same expressions per component, the
// reads and writes of the particle's vectors spread over the group.  For a
// fixed grid height without a wmode-3 wall (the fused pass).  Returns the
// new position's component q (w: the radius).
__device__ __forceinline__ double integrate_component(CParams& p, const DevArrays& a,
                                                      const DevLoop* lp, uint32_t orig, uint32_t q,
                                                      double force_q, double torque_q,
                                                      const double4 x, const double4 vel,
                                                      const double4 angvel, uint32_t r) {
    const uint32_t nxt = cur_set(lp) ^ 1u;
    if (q == 3) {
        a.xs[nxt][r].w = x.w;
        a.vs[nxt][r].w = vel.w;
        a.ws[nxt][r].w = angvel.w;
        return x.w;
    }
    const double t_now = lp->t;
    const int64_t iter = lp->iter;
    int cvs = 0;
    if (t_now >= 10.0 && t_now < 15.0 && lp->velocity_state == 1) cvs = 1;
    else if (t_now >= 15.0 && lp->velocity_state == 2) cvs = -1;
    const double xq = q == 0 ? x.x : q == 1 ? x.y : x.z;
    const double vq = q == 0 ? vel.x : q == 1 ? vel.y : vel.z;
    const double wq = q == 0 ? angvel.x : q == 1 ? angvel.y : angvel.z;
    double* const angpos_c = reinterpret_cast<double*>(a.angpos + orig) + q;
    double* const xyzsum_c = reinterpret_cast<double*>(a.xyzsum + orig) + q;
    double* const acc_c = reinterpret_cast<double*>(a.acc + orig) + q;
    double* const angacc_c = reinterpret_cast<double*>(a.angacc + orig) + q;
    const double angpos_q = *angpos_c;
    double xyzsum_q = *xyzsum_c;
    double acc0 = 0.0, angacc0 = 0.0;
    if (iter != 0) {
        acc0 = *acc_c;
        angacc0 = *angacc_c;
    }
    const double radius = x.w;
    const double dt = p.dt;
    const double m = 4.0 / 3.0 * PI * radius * radius * radius * p.rho;
    double acc = force_q / m + p.g[q];
    double angacc = torque_q * 1.0 / (2.0 / 5.0 * m * radius * radius);
    if (vel.w > 0.0001 && vel.w < 10.0) {
        if (q < 2) acc = 0.0;
        else acc -= p.g[2];
        angacc = 0.0;
    }
    if (vel.w >= 10.0) {
        if (q == 1) acc = 0.0;
        if (q != 1) angacc = 0.0;
    }
    if (vel.w < -0.0001) {
        acc = 0.0;
        angacc = 0.0;
    }
    const double dacc_dt = (acc - acc0) / dt;
    const double dangacc_dt = (angacc - angacc0) / dt;
    double x_new = xq + vq * dt + 0.5 * acc * dt * dt + 1.0 / 6.0 * dacc_dt * dt * dt * dt;
    double vel_new = vq + acc * dt + 0.5 * dacc_dt * dt * dt;
    const double angpos_new =
        angpos_q + wq * dt + 0.5 * angacc * dt * dt + 1.0 / 6.0 * dangacc_dt * dt * dt * dt;
    const double angvel_new = wq + angacc * dt + 0.5 * dangacc_dt * dt * dt;
    xyzsum_q += vq * dt + 0.5 * acc * dt * dt + 1.0 / 6.0 * dacc_dt * dt * dt * dt;
    // Periodic wrap, src/integration.cuh:267-283 (note `> L`, P17).
    if ((p.periodic == 1 && q < 2) || (p.periodic == 2 && q == 0)) {
        if (x_new < p.origo[q]) x_new += p.L[q];
        if (x_new > p.L[q]) x_new -= p.L[q];
    }
    if (q == 0 && vel.w > 5.0 && vel.w < 10.0) {
        if (cvs == 1) vel_new *= 10.0;
        else if (cvs == -1) vel_new /= 10.0;
    }
    *xyzsum_c = xyzsum_q;
    *acc_c = acc;
    *angacc_c = angacc;
    reinterpret_cast<double*>(a.angvel + orig)[q] = angvel_new;
    reinterpret_cast<double*>(a.vel + orig)[q] = vel_new;
    *angpos_c = angpos_new;
    reinterpret_cast<double*>(a.x + orig)[q] = x_new;
    reinterpret_cast<double*>(a.xs[nxt] + r)[q] = x_new;
    reinterpret_cast<double*>(a.vs[nxt] + r)[q] = vel_new;
    reinterpret_cast<double*>(a.ws[nxt] + r)[q] = angvel_new;
    return x_new;
}

// The next step's hash flags of a wavefront, and the particles that changed
// cell listed for the end-of-step re-sort (one atomic per wavefront; the
// list is unordered, the count keeps growing past MV_CAP so that an
// overfull list is recognised).  Every lane of a wavefront that has any
// particle left calls it (f is empty for lanes without one).
__device__ __forceinline__ void publish_next(CParams& p, const DevArrays& a, DevLoop* lp,
                                             const HashFlags& f, uint32_t orig,
                                             uint32_t next_slot) {
    publish_hash(lp, a.glob, next_slot, f);
    const unsigned long long mv = __ballot(f.changed);
    if (mv) {
        const uint32_t lane = threadIdx.x & 63u;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&lp->nmovers, (uint32_t)__popcll(mv));
        base = (uint32_t)__shfl((int)base, 0, 64);
        const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        const uint32_t pos = base + (uint32_t)__popcll(mv & lt);
        if (f.changed && pos < MV_CAP) a.movers[(size_t)p.sim * MV_CAP + pos] = orig;
    }
}

// ---------------------------------------------------------------------------
// topology + interact with LPP lanes per particle (contact models 2 and 3).
// The per-particle work is a chain of dependent gathers (candidate entry ->
// candidate position -> partner state); one lane per particle leaves the
// chip with ~1.5 wavefronts per SIMD at 1e5 particles and nothing to hide
// that latency.  Here a group of LPP consecutive lanes shares a particle:
// the candidates are tested LPP at a time, the contact slots are held LPP-
// strided in registers (slot s on lane s % LPP) during the search, the
// occupied slots of the wavefront's particles are then evaluated one per
// lane of the wavefront, and everything whose order the reference fixes is
// kept in that order: slot registration in scan order, and the
// accumulation of contact contributions in slot order.  Every branch a group
// takes is taken by all of its lanes, so group ballots and shuffles see the
// whole group.

template <int LPP>
struct LaneGroup {
    static_assert(LPP == 4 || LPP == 8 || LPP == 16, "group size");
    static constexpr uint32_t MASK = (1u << LPP) - 1u;
    uint32_t q;     // lane within the group
    uint32_t base;  // wavefront lane of the group's lane 0
    __device__ __forceinline__ LaneGroup() {
        const uint32_t l = __lane_id();
        q = l % LPP;
        base = l - q;
    }
    // Bit i = predicate of group lane i.
    __device__ __forceinline__ uint32_t bits(bool v) const {
        return (uint32_t)(__ballot(v) >> base) & MASK;
    }
    __device__ __forceinline__ uint32_t from(uint32_t v, uint32_t src) const {
        return (uint32_t)__shfl((int)v, (int)(base + src), 64);
    }
    __device__ __forceinline__ double from(double v, uint32_t src) const {
        return __shfl(v, (int)(base + src), 64);
    }
    __device__ __forceinline__ d3 from(d3 v, uint32_t src) const {
        return mk3(from(v.x, src), from(v.y, src), from(v.z, src));
    }
    // Group lane T's value, T fixed at compile time: for groups of four (a
    // DPP quad) one quad_perm move per dword, with no LDS round trip.
    template <int T>
    __device__ __forceinline__ uint32_t bc(uint32_t v) const {
        if constexpr (LPP == 4) {
            constexpr int ctrl = T | (T << 2) | (T << 4) | (T << 6);
            return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, ctrl, 0xf, 0xf, false);
        } else {
            return from(v, (uint32_t)T);
        }
    }
    template <int T>
    __device__ __forceinline__ double bc(double v) const {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = bc<T>((uint32_t)u), hi = bc<T>((uint32_t)(u >> 32));
        return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
    }
    template <int T>
    __device__ __forceinline__ d3 bc(d3 v) const {
        return mk3(bc<T>(v.x), bc<T>(v.y), bc<T>(v.z));
    }
};

// Copy of n contiguous words (n = 2, 4 or 8; both sides 8-byte aligned,
// 16-byte aligned for n >= 4).
template <int N>
__device__ __forceinline__ void copy_words(uint32_t* dst, const uint32_t* src) {
    if constexpr (N % 4 == 0) {
#pragma unroll
        for (int k = 0; k < N / 4; ++k)
            reinterpret_cast<uint4*>(dst)[k] = reinterpret_cast<const uint4*>(src)[k];
    } else {
        static_assert(N == 2, "word count");
        *reinterpret_cast<uint2*>(dst) = *reinterpret_cast<const uint2*>(src);
    }
}

// Orders the wavefront's LDS accesses around it (the lanes of a group
// exchange row words through LDS; a group never spans wavefronts).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}


// The contact-slot policy (src/contactsearch.cuh:265-299) over the slot
// registers of a group: the first slot holding the partner, else the last
// empty slot, else slot nc (whose write is suppressed, as in the one-lane
// kernel; `clobber` is then set).  Returns true when nc exceeds NC, which
// ends the current cell.
template <int LPP>
__device__ __forceinline__ bool group_register(const LaneGroup<LPP>& g, uint32_t (&slot)[NC / LPP],
                                               uint32_t word, int& nc, bool& clobber,
                                               DevLoop* loop) {
    constexpr int SPL = NC / LPP;
    const uint32_t ib = word & CW_IDX_MASK;
    uint32_t match = 0, empty = 0;
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
        const uint32_t cidx = slot[j] & CW_IDX_MASK;
        match |= g.bits(cidx == ib) << (LPP * j);
        empty |= g.bits(cidx == EMPTY) << (LPP * j);
    }
    int cpos = nc;
    if (match) cpos = __ffs(match) - 1;
    else if (empty) cpos = 31 - __clz((int)empty);
    else clobber = true;
    if (cpos < NC && (uint32_t)cpos % LPP == g.q) {
#pragma unroll
        for (int j = 0; j < SPL; ++j)
            if (j == cpos / LPP) slot[j] = word;
    }
    ++nc;
    if (nc > NC) {
        if (g.q == 0) atomicOr(&loop->error_bits, (uint32_t)DFLAG_CONTACT_OVERFLOW);
        return true;
    }
    return false;
}

// Overlap tests of a lane's list entries (entry g.q + LPP i in ev[i]) whose
// bit is set in `want`, in chunks of TCHUNK: each chunk's partner positions
// are loaded together (an entry not tested loads the particle's own
// position) before any of them is used, so that a chunk costs one memory
// round trip.  The second chunk runs only when some lane of the wavefront
// lists entries there (`tail`).  Returns bit i set where entry i overlaps:
// the same test as the full scan's, len32 - (r_a + r_b) < 0 after the
// conservative reject dd >= S^2 (1 + 2^-16).
constexpr int TCHUNK = 4;
template <int LPP>
__device__ __forceinline__ uint32_t test_entries(CParams& p, const SortedReader& S,
                                                 const double4 xa, uint32_t ia,
                                                 const uint2 (&ev)[NV / LPP], uint32_t want,
                                                 bool tail) {
    constexpr int RMAX = NV / LPP;
    uint32_t hits = 0;
#pragma unroll
    for (int c = 0; c < RMAX; c += TCHUNK) {
        if (c > 0 && !tail) break;
        // The chunk's indices first, then (stale copies only) their original
        // indices together, then the positions together: one wait per step
        // of the chain, not one per candidate.
        uint32_t ib[TCHUNK];
#pragma unroll
        for (int k = 0; k < TCHUNK; ++k)
            ib[k] = (c + k < RMAX && ((want >> (c + k)) & 1u)) ? (ev[c + k].x & VL_IB_MASK) : ia;
        if (S.stale) {
#pragma unroll
            for (int k = 0; k < TCHUNK; ++k) ib[k] = S.perm[ib[k]];
        }
        double4 xb[TCHUNK];
#pragma unroll
        for (int k = 0; k < TCHUNK; ++k)
            if (c + k < RMAX) xb[k] = S.x[ib[k]];
#pragma unroll
        for (int k = 0; k < TCHUNK; ++k) {
            if (c + k >= RMAX) continue;
            const int i = c + k;
            const double rb = xb[k].w;
            d3 xab = mk3(xa.x - xb[k].x, xa.y - xb[k].y, xa.z - xb[k].z);
            xab += shift_of(p, word_code(ev[i].y));
            const double dd = dot(xab, xab);
            const double S = xa.w + rb;
            if (((want >> i) & 1u) && !(dd >= S * S * 1.0000152587890625)) {
                const double len = sqrt32(dd);
                if (len - S < 0.0) hits |= 1u << i;
            }
        }
    }
    return hits;
}

// OR of a value over the lanes of a group.
template <int LPP>
__device__ __forceinline__ uint32_t group_or(uint32_t v) {
#pragma unroll
    for (int m = 1; m < LPP; m <<= 1) v |= (uint32_t)__shfl_xor((int)v, m, 64);
    return v;
}

// The common case of a round of list entries: every overlapping entry of
// the round (bit t of hm: group lane t's entry e) is a partner that exactly
// one slot of the row holds, the cell is not done and no overflow is
// possible.  The slot policy then only rewrites each matched slot with the
// entry's word (its image code may have changed) in entry order, and counts
// nc up: done here at once, and true returned.  Otherwise nothing changes
// and false is returned (the entries are then registered one by one).
// Groups of four only (DPP quad broadcasts).
template <int LPP>
__device__ __forceinline__ bool group_register_known(const LaneGroup<LPP>& g,
                                                     uint32_t (&slot)[NC / LPP], uint2 e,
                                                     uint32_t hm, bool cell_done, int& nc,
                                                     uint32_t& last_cell) {
    if constexpr (LPP != 4) {
        return false;
    } else {
        constexpr int SPL = NC / LPP;
        const int nh = __popc(hm);
        if (cell_done || nc + nh > NC) return false;
        uint32_t idx[4];
        idx[0] = g.template bc<0>(e.y) & CW_IDX_MASK;
        idx[1] = g.template bc<1>(e.y) & CW_IDX_MASK;
        idx[2] = g.template bc<2>(e.y) & CW_IDX_MASK;
        idx[3] = g.template bc<3>(e.y) & CW_IDX_MASK;
        // mm bit t * SPL + j: this lane's slot j holds hit t's partner.
        uint32_t mm = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int j = 0; j < SPL; ++j)
                if (((hm >> t) & 1u) && (slot[j] & CW_IDX_MASK) == idx[t])
                    mm |= 1u << (t * SPL + j);
        // Byte t: the group's number of slots holding hit t's partner.
        uint32_t cnt = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
            cnt += (uint32_t)__popc((mm >> (t * SPL)) & ((1u << SPL) - 1u)) << (8 * t);
        cnt += (uint32_t)__shfl_xor((int)cnt, 1, 64);
        cnt += (uint32_t)__shfl_xor((int)cnt, 2, 64);
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (((cnt >> (8 * t)) & 0xffu) != ((hm >> t) & 1u)) return false;
        uint32_t w[4];
        w[0] = g.template bc<0>(e.y);
        w[1] = g.template bc<1>(e.y);
        w[2] = g.template bc<2>(e.y);
        w[3] = g.template bc<3>(e.y);
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int j = 0; j < SPL; ++j)
                if ((mm >> (t * SPL + j)) & 1u) slot[j] = w[t];
        nc += nh;
        last_cell = g.from(e.x, 31u - (uint32_t)__clz((int)hm)) >> VL_CELL_SHIFT;
        return true;
    }
}

// Four wavefronts per SIMD (128 VGPRs): the compiler's own allocation takes
// 131 and three; capping it spills 12 bytes per lane and measured 51.3
// against 54.9 us on the 1e5-particle sheared bed.
// WG work-items per workgroup (WG / LPP particles, within one 256-particle
// chunk).
template <int CM, int LPP, bool FUSED, int WG>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(4))) void k_contact_g(
    const DevArrays a) {
    static_assert(CM == 2 || CM == 3, "history models only");
    static_assert(WG % LPP == 0 && BLOCK % (WG / LPP) == 0, "workgroups inside a chunk");
    constexpr int SPL = NC / LPP;       // slots per lane
    constexpr int PPB = WG / LPP;       // particles per workgroup
    const LaneGroup<LPP> g;
    const uint32_t blk = xcd_block(a, blockIdx.x, gridDim.x);
    const uint32_t ia = blk * PPB + threadIdx.x / LPP;
    // A workgroup's particles lie in one 256-particle chunk (computed from
    // the workgroup index alone, so that the parameter loads are scalar).
    CParams& p = chunk_params(a, blk * PPB / BLOCK);
    // The interaction's staging (per wavefront): each group's number of
    // occupied slots, whether one of its contacts ended, the contact table
    // (group and slot of each contact), the particles' state and index,
    // and one round's contributions (6 doubles per lane).
    constexpr int G = 64 / LPP;   // groups per wavefront
    constexpr int NWV = WG / 64;  // wavefronts per workgroup
    __shared__ __align__(16) uint8_t s_zc[NWV][G];
    __shared__ uint8_t s_end[NWV][G];
    __shared__ uint16_t s_tab[NWV][G * NC];
    __shared__ double4 s_pa[NWV][3][G];
    __shared__ uint32_t s_po[NWV][G];
    __shared__ double s_con[NWV][6][64];
    const uint32_t wv = threadIdx.x / 64u;
    const uint32_t gw = (threadIdx.x & 63u) / LPP;  // group within the wavefront
    // Groups past the end count no contacts.
    if (g.q == 0) s_zc[wv][gw] = 0;
    if (ia - p.base >= p.np) return;  // whole groups leave together
    DevLoop* lp = a.loop + p.sim;
    // The loads that depend only on the particle's sorted index go out
    // together: its original index, its list length and hit mask, and its
    // position (from the current sorted copy, read whether or not the
    // copies are stale, and replaced by the original array's when they are,
    // so that the common case does not wait for the original index).
    const uint32_t ia_orig = a.perm[ia];
    const SortedReader S(a, lp);
    const bool rebuild =
        !p.vl_enabled || lp->rebuild[step_slot(lp)] != 0u || lp->force_rebuild != 0u;
    const uint2 vmeta = rebuild ? make_uint2(0u, 0u) : a.vcount[ia];
    double4 xa = a.xs[cur_set(lp)][ia];
    if (S.stale) xa = a.x[ia_orig];
    const double ra = xa.w;
    const double Lz = grid_Lz(p, a);
    // The row passes through LDS: read and written as contiguous words per
    // lane, held LPP-strided in registers during the search, and back in
    // LDS for the interaction.
    __shared__ uint32_t s_slot[PPB * NC];
    uint32_t* gs = s_slot + (threadIdx.x / LPP) * NC;
    uint32_t* const crow_a = a.crow + (size_t)ia_orig * NC;
    const uint32_t vcnt = vmeta.x == VL_OVERFLOW ? VL_OVERFLOW : (vmeta.x & VC_COUNT_MASK);
    const bool hp_valid = vmeta.x != VL_OVERFLOW && (vmeta.x & VC_HP_VALID) != 0u;
    // The list entries (list path) are requested before the row, so that
    // both are in flight together.
    constexpr int RMAX = NV / LPP;
    const bool list_path = !rebuild && vcnt != VL_OVERFLOW;
    uint2 ev[RMAX];
    {
        const uint2* vrow = a.vlist + (size_t)ia * NV;
#pragma unroll
        for (int i = 0; i < RMAX; ++i) {
            ev[i] = make_uint2(0u, 0u);
            if (list_path && g.q + LPP * i < vcnt) ev[i] = vrow[g.q + LPP * i];
        }
    }
    copy_words<SPL>(gs + g.q * SPL, crow_a + g.q * SPL);
    wave_sync();
    // Slot registers; after a re-sort the words are mapped to the new order.
    uint32_t slot[SPL];
    const bool remap = lp->order_changed != 0u;
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
        slot[j] = gs[g.q + LPP * j];
        if (remap) slot[j] = word_to_current(a, slot[j]);
    }

    // topology: the candidate sequence of the 27-cell scan, LPP at a time.
    int nc = 0;
    const uint32_t below = (1u << g.q) - 1u;
    // In the list path: whether the row after this step holds exactly the
    // partners of the overlapping entries (hp_new, VC_HP_VALID).
    bool quiet = false;
    if (rebuild || vcnt == VL_OVERFLOW) {
        const bool build = p.vl_enabled && rebuild;
        uint32_t nl = 0;
        bool clobber = false;
        uint2* vrow = a.vlist + (size_t)ia * NV;
        const int gx = (int)floor((xa.x - p.origo[0]) / (p.L[0] / p.num[0]));
        const int gy = (int)floor((xa.y - p.origo[1]) / (p.L[1] / p.num[1]));
        const int gz = (int)floor((xa.z - p.origo[2]) / (Lz / p.num[2]));
        for (int dz = -1; dz < 2; ++dz)
            for (int dy = -1; dy < 2; ++dy)
                for (int dx = -1; dx < 2; ++dx) {
                    int tx = gx + dx, ty = gy + dy;
                    const int tz = gz + dz;
                    const int code = find_dist_mod(p, tx, ty, tz);
                    if (code < 0) continue;
                    const uint32_t cell = p.cell_base + (uint32_t)tx + (uint32_t)ty * p.num[0] +
                                          (p.num[0] * p.num[1]) * (uint32_t)tz;
                    const uint32_t start = a.cell_start[cell];
                    if (start == 0xffffffffu) continue;
                    const uint32_t end = a.cell_end[cell];
                    const d3 dm = shift_of(p, (uint32_t)code);
                    const uint32_t celloff = (uint32_t)((dz + 1) * 9 + (dy + 1) * 3 + (dx + 1));
                    bool cell_done = false;
                    for (uint32_t r = start; r < end; r += LPP) {
                        const uint32_t ib = r + g.q;
                        bool hit = false, near = false;
                        uint32_t word = 0;
                        if (ib < end && ib != ia) {
                            const double4 xb = S.x[S.at(ib)];
                            const double rb = xb.w;
                            d3 xab = mk3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z);
                            xab += dm;
                            const double dd = dot(xab, xab);
                            const double Sab = ra + rb;
                            if (build) {
                                const double Sv = Sab + p.vl_skin;
                                near = dd < Sv * Sv;
                            }
                            // Conservative reject: len32 > S whenever
                            // dd >= S^2 (1 + 2^-16).
                            if (!(dd >= Sab * Sab * 1.0000152587890625)) {
                                const double len = sqrt32(dd);
                                hit = len - Sab < 0.0;
                            }
                            if (hit || near) word = pack_word(ib, (uint32_t)code);
                        }
                        if (build) {
                            const uint32_t nm = g.bits(near);
                            const uint32_t pos = nl + __popc(nm & below);
                            if (near && pos < (uint32_t)NV)
                                vrow[pos] = make_uint2(ib | (celloff << VL_CELL_SHIFT), word);
                            nl += __popc(nm);
                        }
                        uint32_t hm = g.bits(hit);
                        while (hm && !cell_done) {
                            const uint32_t b = __ffs(hm) - 1;
                            hm &= hm - 1u;
                            cell_done = group_register(g, slot, g.from(word, b), nc, clobber, lp);
                        }
                    }
                }
        if (build && g.q == 0) {
            // A new list starts without a hit mask.
            a.vcount[ia] = make_uint2(nl <= p.vl_cap ? nl : VL_OVERFLOW, 0u);
            a.x_build[ia_orig] = xa;
        }
    } else {
        // Every listed candidate is tested first (lane q takes entries q,
        // q + LPP, ...; all their loads are in flight together).  With a
        // valid hit mask only the entries that held no contact are tested:
        // when none of them overlaps, the slot policy would only re-register
        // each held partner in its own slot with the word it holds, so the
        // row stays as it is, and a held contact that ended is emptied by
        // the interaction below exactly as without registration.  Otherwise
        // the held entries are tested too and the overlapping entries
        // register in list (= scan) order.
        const uint32_t held = hp_valid ? (vmeta.y >> (RMAX * g.q)) & ((1u << RMAX) - 1u) : 0u;
        // The lane's listed entries as a bit mask, and whether any lane of
        // the wavefront has entries beyond the first chunk of tests.
        const uint32_t listed =
            vcnt > g.q ? ((1u << ((vcnt - g.q + LPP - 1) / LPP)) - 1u) & ((1u << RMAX) - 1u) : 0u;
        const bool tail = __ballot(listed >> TCHUNK) != 0ull;
        // Entries not held: their tests decide whether registration is needed.
        uint32_t hits = test_entries<LPP>(p, S, xa, ia, ev, listed & ~held, tail);
        quiet = hp_valid && g.bits(hits != 0u) == 0u;
        if (!quiet) {
            hits |= test_entries<LPP>(p, S, xa, ia, ev, listed & held, tail);
            const uint32_t hits0 = hits;
            // Only overlapping candidates register; the cells of the others
            // cannot matter, as each cell's entries are contiguous.  The
            // entries are shifted down a round at a time (never indexed).
            uint32_t last_cell = 0xffffffffu;
            bool cell_done = false, clobber = false, overflow = false;
            for (uint32_t r = 0; r < vcnt; r += LPP) {
                uint32_t hm = g.bits(hits & 1u);
                const uint2 e = ev[0];
#pragma unroll
                for (int i = 0; i + 1 < RMAX; ++i) ev[i] = ev[i + 1];
                hits >>= 1;
                if (hm && group_register_known<LPP>(g, slot, e, hm, cell_done, nc, last_cell))
                    continue;
                while (hm) {
                    const uint32_t b = __ffs(hm) - 1;
                    hm &= hm - 1u;
                    const uint32_t celloff = g.from(e.x, b) >> VL_CELL_SHIFT;
                    if (celloff != last_cell) {
                        last_cell = celloff;
                        cell_done = false;
                    }
                    if (cell_done) {
                        overflow = true;
                        continue;
                    }
                    cell_done = group_register(g, slot, g.from(e.y, b), nc, clobber, lp);
                    overflow = overflow || cell_done;
                }
            }
            // The row now holds exactly the overlapping entries' partners
            // unless a registration found neither its partner nor an empty
            // slot, or the count passed NC; the interaction keeps exactly
            // those (it applies the same overlap test).
            const uint32_t hp_new = group_or<LPP>(hits0 << (RMAX * g.q));
            const bool exact = g.bits(clobber || overflow) == 0u;
            if (g.q == 0)
                a.vcount[ia] = make_uint2(vcnt | (exact ? VC_HP_VALID : 0u), hp_new);
        }
    }

    // The particle's velocities are loaded only now: the candidate tests
    // above hold many positions in flight, and registers are what bound the
    // number of wavefronts per SIMD here.
    const double4 vel_a = S.v[S.stale ? ia_orig : ia];
    const double4 av_a = S.w[S.stale ? ia_orig : ia];
    const d3 angvel_a = xyz(av_a);
    tape_start(a, p, ia, ia_orig, xa, vel_a, av_a, g.q == 0);

    // interact (src/contactsearch.cuh:484-559).  The occupied slots of the
    // wavefront's groups are numbered group by group, each group's in
    // ascending slot order, and evaluated one per lane of the wavefront at
    // a time, so that no particle's contacts wait for those of the
    // wavefront's most connected particle.  Each contact's contribution
    // goes to LDS; each group then adds its contacts' contributions in slot
    // order: lane 0 the force, lane 1 the torque, lane 2 p and the energy
    // rates.  An absent energy addition is stored as +0.0: the sums start
    // at +0.0 and so are never -0.0, and x + 0.0 = x for every other x.
    const bool zero_empty = lp->zero_empty != 0u;
    uint32_t occ = 0;
    wave_sync();
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
        const bool o = (slot[j] & CW_IDX_MASK) != EMPTY;
        occ |= g.bits(o) << (LPP * j);
        gs[g.q + LPP * j] = slot[j];
        if (!o && zero_empty)
            a.delta_t[(size_t)ia_orig * NC + g.q + LPP * j] = make_double4(0.0, 0.0, 0.0, 0.0);
    }
    const uint32_t zg = __popc(occ);
    if (g.q == 0) {
        s_zc[wv][gw] = (uint8_t)zg;
        s_end[wv][gw] = 0;
        s_pa[wv][0][gw] = xa;
    }
    if (g.q == 1) s_pa[wv][1][gw] = vel_a;
    if (g.q == 2) s_pa[wv][2][gw] = av_a;
    if (g.q == 3) s_po[wv][gw] = ia_orig;
    wave_sync();
    // The group's first contact number and the wavefront's contact count
    // (byte sums of the counts, each at most NC).
    uint32_t first = 0, total = 0;
    {
        const uint32_t* zw = reinterpret_cast<const uint32_t*>(s_zc[wv]);
#pragma unroll
        for (int i = 0; i < G / 4; ++i) {
            const uint32_t w = zw[i];
            const int below_w = (int)gw - 4 * i;
            const uint32_t m = below_w <= 0 ? 0u
                               : below_w >= 4 ? 0xffffffffu
                                              : (1u << (8 * below_w)) - 1u;
            first += ((w & m) * 0x01010101u) >> 24;
            total += (w * 0x01010101u) >> 24;
        }
    }
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
        const uint32_t sl = g.q + LPP * j;
        if ((occ >> sl) & 1u)
            s_tab[wv][first + __popc(occ & ((1u << sl) - 1u))] = (uint16_t)((gw << 5) | sl);
    }
    wave_sync();
    // The wavefront's lanes still running (groups past the end have left)
    // take the contacts in turn, by their rank among the running lanes.
    const uint64_t act = __ballot(1);
    const uint32_t nact = (uint32_t)__popcll(act);
    const uint32_t rk = __builtin_amdgcn_mbcnt_hi((uint32_t)(act >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
    double4 acc = make_double4(0.0, 0.0, 0.0, 0.0);
    for (uint32_t r0 = 0; r0 < total; r0 += nact) {
        const uint32_t c = r0 + rk;
        Contrib C;
        contrib_init(C);
        C.F = mk3(0.0, 0.0, 0.0);
        C.T = mk3(0.0, 0.0, 0.0);
        C.p = 0.0;
        C.ev0 = 0.0;
        C.ev1 = 0.0;
        C.es = 0.0;
        if (c < total) {
            const uint32_t e = s_tab[wv][c];
            const uint32_t h = e >> 5, sidx = e & 31u;
            uint32_t* const hs = s_slot + (wv * G + h) * NC;
            const double4 xc = s_pa[wv][0][h];
            const double4 vel_c = s_pa[wv][1][h];
            const double4 av_c = s_pa[wv][2][h];
            const uint32_t io = s_po[wv][h];
            const uint32_t wrd = hs[sidx];
            const size_t mp = (size_t)io * NC + sidx;
            const uint32_t rb_s = wrd & CW_IDX_MASK;
            const d3 dm = shift_of(p, word_code(wrd));
            // A slot's contact has almost always persisted: its partner's
            // velocities and the displacement are loaded with its position.
            const uint32_t rb_j = S.at(rb_s);
            const double4 xb = S.x[rb_j];
            const double4 vel_b = S.v[rb_j];
            const double4 av_b = S.w[rb_j];
            const double4 dt0 = a.delta_t[mp];
            const double rc = xc.w;
            const double rb = xb.w;
            const d3 xab = mk3(xc.x - xb.x + dm.x, xc.y - xb.y + dm.y, xc.z - xb.z + dm.z);
            const double len = length(xab);
            const double dn = len - (rc + rb);
            if (dn < 0.0) {
                if (a.tape_dt0) {
                    a.tape_dt0[mp] = dt0;
                    a.tape_srt[mp] = rb_s;
                }
                d3 nd;
                if constexpr (CM == 2)
                    nd = contact_linear(p, C, vel_c, xyz(av_c), vel_b, xyz(av_b), rc, rb, xab,
                                        len, dn, dt0);
                else
                    nd = contact_hertz(p, C, vel_c, xyz(av_c), vel_b, xyz(av_b), rc, rb, xab,
                                       len, dn, dt0);
                a.delta_t[mp] = make_double4(nd.x, nd.y, nd.z, 0.0);
            } else {
                // The contact ended: its slot empties (src/contactsearch.cuh:540-548).
                hs[sidx] = EMPTY;
                a.delta_t[mp] = make_double4(0.0, 0.0, 0.0, 0.0);
                s_end[wv][h] = 1;
            }
        }
        // The group's contacts of this round, in slot order, in two passes
        // through one buffer: force and torque (lanes 0 and 1, x, y, z),
        // then p, ev_dot += ev0, es_dot and ev_dot += ev1 (lane 2).
        const uint32_t lo = first > r0 ? first : r0;
        const uint32_t hi = first + zg < r0 + nact ? first + zg : r0 + nact;
        s_con[wv][0][rk] = C.F.x;
        s_con[wv][1][rk] = C.F.y;
        s_con[wv][2][rk] = C.F.z;
        s_con[wv][3][rk] = C.T.x;
        s_con[wv][4][rk] = C.T.y;
        s_con[wv][5][rk] = C.T.z;
        wave_sync();
        if (g.q < 2) {
            for (uint32_t cc = lo; cc < hi; ++cc) {
                const uint32_t l = cc - r0;
                acc.x += s_con[wv][3 * g.q][l];
                acc.y += s_con[wv][3 * g.q + 1][l];
                acc.z += s_con[wv][3 * g.q + 2][l];
            }
        }
        wave_sync();
        s_con[wv][0][rk] = C.p;
        s_con[wv][1][rk] = C.nev > 0 ? C.ev0 : 0.0;
        s_con[wv][2][rk] = C.nes > 0 ? C.es : 0.0;
        s_con[wv][3][rk] = C.nev > 1 ? C.ev1 : 0.0;
        wave_sync();
        if (g.q == 2) {
            for (uint32_t cc = lo; cc < hi; ++cc) {
                const uint32_t l = cc - r0;
                acc.x += s_con[wv][0][l];
                acc.y += s_con[wv][1][l];
                acc.z += s_con[wv][2][l];
                acc.y += s_con[wv][3][l];
            }
        }
        wave_sync();
    }
    // Each lane's sums in the layout of Acc (lane 0: F, lane 1: T, lane 2:
    // p, ev_dot, es_dot); the walls add to every field alike, and each lane
    // writes its part of the output record.
    Acc A;
    A.F = mk3(acc.x, acc.y, acc.z);
    A.T = mk3(acc.x, acc.y, acc.z);
    A.p = acc.x;
    A.ev_dot = acc.y;
    A.es_dot = acc.z;
    const bool ended = s_end[wv][gw] != 0;
    // The particle's state again, from LDS (not held in registers through
    // the interaction).
    const double4 xa_o = s_pa[wv][0][gw];
    const double4 vel_o = s_pa[wv][1][gw];
    const double4 av_o = s_pa[wv][2][gw];
    // Without registration the hit mask stays true until a contact ends,
    // and the row changes only where a contact ended or the order changed.
    const bool any_ended = ended;
    if (quiet && any_ended && g.q == 0) a.vcount[ia] = make_uint2(vcnt, 0u);
    wave_sync();
    if (!quiet || any_ended || remap)
        copy_words<SPL>(a.crow + (size_t)ia_orig * NC + g.q * SPL, gs + g.q * SPL);
    if (a.tape_rows) {
#pragma unroll
        for (int k = 0; k < SPL; ++k)
            a.tape_rows[(size_t)ia_orig * NC + g.q * SPL + k] = word_to_orig(a, gs[g.q * SPL + k]);
    }
    walls_and_outputs(p, a, A, xa_o, vel_o, av_o, Lz, ia_orig, g.q < 4 ? (int)g.q : -1);
    if constexpr (FUSED) {
        // The particle's force and torque are final (its own row and the
        // walls): integrate it now (src/integration.cuh:14-305), from the
        // step's sorted copies, into the next step's set, and hash it for
        // the next step.  (Fixed grid height, no bonds, no wmode-3 wall.)
        static_assert(!FUSED || LPP == 4, "one vector component per lane");
        // Lane 0 holds the force, lane 1 the torque, lane 2 the energy
        // rates: lane q takes component q of each.
        if (g.q == 2) {
            a.es[ia_orig] = a.es[ia_orig] + A.es_dot * p.dt;
            a.ev[ia_orig] = a.ev[ia_orig] + A.ev_dot * p.dt;
        }
        const double fx = g.template bc<0>(A.F.x), fy = g.template bc<0>(A.F.y),
                     fz = g.template bc<0>(A.F.z);
        const double tx = g.template bc<1>(A.T.x), ty = g.template bc<1>(A.T.y),
                     tz = g.template bc<1>(A.T.z);
        const double fq = g.q == 0 ? fx : g.q == 1 ? fy : fz;
        const double tq = g.q == 0 ? tx : g.q == 1 ? ty : tz;
        const double xn = integrate_component(p, a, lp, ia_orig, g.q, fq, tq, xa_o, vel_o, av_o, ia);
        const double4 x_new = make_double4(g.template bc<0>(xn), g.template bc<1>(xn),
                                           g.template bc<2>(xn), g.template bc<3>(xn));
        HashFlags f;
        if (g.q == 0) hash_particle(p, a, ia_orig, p.L[2], x_new, f);
        publish_next(p, a, lp, f, ia_orig, step_slot(lp) ^ 1u);
    }
}

// ---------------------------------------------------------------------------
// Block tree sum of the reference's `summation` kernel (one element per
// work-item, then the halving tree, src/integration.cuh:309-341).

// The halving steps of stride 128 and 64 cross wavefronts and go through
// LDS; those of stride 32..1 stay in wavefront 0 and use shuffles.  Every
// element receives the same additions in the same order as in the LDS tree.
// The sum is valid on work-item 0.
__device__ __forceinline__ double block_tree_sum(double v, double* cache) {
    static_assert(BLOCK == 256, "tree of 256");
    double temp = 0.0;
    temp += v;
    cache[threadIdx.x] = temp;
    __syncthreads();
    if (threadIdx.x < 128) cache[threadIdx.x] += cache[threadIdx.x + 128];
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x < 64) {
        s = cache[threadIdx.x] + cache[threadIdx.x + 64];
#pragma unroll
        for (int i = 32; i > 0; i >>= 1) {
            const double o = __shfl_down(s, (unsigned)i, 64);
            if ((int)threadIdx.x < i) s += o;
        }
    }
    __syncthreads();  // cache may be reused by the caller
    return s;
}

// The serial left-to-right sum of n partials (integrateWalls'
// `for i < blocksPerGrid: sum += partial[i]`), with the partials staged
// through LDS by the whole workgroup.  The result is valid in thread 0.
__device__ __forceinline__ double serial_sum(const double* __restrict__ part, uint32_t n,
                                             double* stage) {
    constexpr uint32_t CH = 2048;
    double sum = 0.0;
    for (uint32_t base = 0; base < n; base += CH) {
        const uint32_t cnt = (n - base) < CH ? (n - base) : CH;
        for (uint32_t j = threadIdx.x; j < cnt; j += BLOCK) stage[j] = part[base + j];
        __syncthreads();
        if (threadIdx.x == 0) {
            // Sixteen LDS reads in flight, then the dependent adds in order:
            // the chain costs one LDS round trip per sixteen partials.
            uint32_t j = 0;
            for (; j + 16 <= cnt; j += 16) {
                double v[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) v[k] = stage[j + k];
#pragma unroll
                for (int k = 0; k < 16; ++k) sum += v[k];
            }
            for (; j < cnt; ++j) sum += stage[j];
        }
        __syncthreads();
    }
    return sum;
}

// findShearStressOnFixedMovingParticles + summation (src/integration.cuh:
// 457-484, 309-341): per-particle stress and the 256-item block partials.
// Partials cross to the serial sum through a kernel boundary, not through
// an in-kernel arrival counter: a device-scope release in every workgroup
// writes back each XCD's L2 and costs more than the launch it saves.
__global__ __launch_bounds__(BLOCK) void k_shear_stress(const DevArrays a) {
    __shared__ double cache[BLOCK];
    CParams& p = chunk_params(a, blockIdx.x);
    const uint32_t idx = blockIdx.x * BLOCK + threadIdx.x;
    const double Lz = grid_Lz(p, a);
    double v = 0.0;
    bool in = idx - p.base < p.np;
    if (in) {
        const double z = a.x[idx].z;
        const double fixvel = a.vel[idx].w;
        const double force_x = a.out[(size_t)idx * 4 + OUT_F].x;
        double f_x = 0.0;
        if (fixvel > 0.0 && z > Lz * 0.5) f_x = force_x;
        v = f_x / (p.L[0] * p.L[1]);
        a.tau_pp[idx] = v;
    }
    const double s = block_tree_sum(in ? v : 0.0, cache);
    if (threadIdx.x == 0) a.tau_partial[blockIdx.x] = s;
}

// The serial sum of the stress partials that integrate reads as tau_eff_x
// (src/integration.cuh:64-68); one workgroup per simulation.
__global__ __launch_bounds__(BLOCK) void k_tau_sum(const DevArrays a) {
    __shared__ double stage[2048];
    CParams& p = sim_params(a, blockIdx.x);
    const double tau = serial_sum(a.tau_partial + p.chunk_base, p.nblocks, stage);
    if (threadIdx.x == 0) a.tau_eff[p.sim] = tau;
}

// ---------------------------------------------------------------------------
// integrate + the next step's hash + the wall force block partials.  The
// grid is the summation kernel's (256 work-items per block over np, so the
// partials are the reference's, src/integration.cuh:309-341).  With a fixed
// grid height the cell keys of step k + 1 depend only on the positions this
// pass writes, so they are computed here into flag slot (k + 1) & 1 and the
// next step needs no hash launch; with an adaptive grid (L[2] follows the
// top wall, which moves after this pass) hash_next is 0.

// Four wavefronts per SIMD (104 VGPRs, no spills).  Five would need 96 and
// spill (with the per-wall partial sums), measured 5-7% slower both on 1e5
// particles and on a 1.45e6-particle batch.
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_integrate(
    const DevArrays a, int hash_next) {
    __shared__ double cache[BLOCK];
    CParams& p = chunk_params(a, blockIdx.x);
    DevLoop* lp = a.loop + p.sim;
    const uint32_t idx = blockIdx.x * BLOCK + threadIdx.x;
    const bool in = idx - p.base < p.np;
    const bool order_changed = lp->order_changed != 0u;
    const uint32_t next_slot = step_slot(lp) ^ 1u;
    HashFlags f;
    double wall0 = 0.0;  // the particle's force on wall 0 (record)
    if (in) {
        // Every load of the particle's state is issued before the first
        // store (a load after a store to memory that may alias it waits for
        // the store): es and ev are advanced (src/contactsearch.cuh:716-717)
        // from values read here and stored after the integration.
        const double4* rec = a.out + (size_t)idx * 4;
        const double4 e = rec[OUT_E];
        const double es0 = a.es[idx], ev0 = a.ev[idx];
        uint32_t key_prev = 0;
        double4 xb = make_double4(0.0, 0.0, 0.0, 0.0);
        if (hash_next) hash_inputs(p, a, idx, key_prev, xb);
        const double4 x_new = integrate_particle(p, a, idx, rec[OUT_F], rec[OUT_T], a.x[idx],
                                                 a.vel[idx], a.angvel[idx], a.rank[idx]);
        a.es[idx] = es0 + e.x * p.dt;
        a.ev[idx] = ev0 + e.y * p.dt;
        wall0 = e.w;
        if (hash_next) hash_particle_from(p, a, idx, p.L[2], x_new, key_prev, xb, f);
    }
    // The contact pass has mapped its rows to the current order.
    if (order_changed) a.perm_rows[idx] = a.perm[idx];
    if (hash_next) publish_next(p, a, lp, f, idx, next_slot);
    // Wall-force block partials: wall 0's, and with per-wall sums every
    // wall's (wall-major, nchunk per wall).
    const uint32_t nsum = p.nw == 0 ? 0u : (p.wall_sums ? p.nw : 1u);
    for (uint32_t w = 0; w < nsum; ++w) {
        const double v = in ? (w == 0 ? wall0 : wall_force_pp(a, idx, w)) : 0.0;
        const double s = block_tree_sum(v, cache);
        if (threadIdx.x == 0) a.partial[(size_t)w * a.nchunk + blockIdx.x] = s;
    }
}

// The integration with one vector component per work-item (four per
// particle, original order): integrate_component's expressions, the same
// as k_integrate's per component, with each particle's vectors read and
// written by its four work-items together.  The next step's hash is taken
// on component lane 0 from the position the quad assembles by DPP.  The
// wall-force block partials follow in k_wall_partials.  For a fixed grid
// height without a wmode-3 wall.
__global__ __launch_bounds__(BLOCK) void k_integrate_q(const DevArrays a) {
    constexpr uint32_t PPB = BLOCK / 4;  // particles per workgroup
    const LaneGroup<4> g;
    const uint32_t idx = blockIdx.x * PPB + threadIdx.x / 4;
    CParams& p = chunk_params(a, blockIdx.x / (BLOCK / PPB));
    DevLoop* lp = a.loop + p.sim;
    const bool in = idx - p.base < p.np;  // the same for a particle's four lanes
    // The contact pass has mapped its rows to the current order.
    if (lp->order_changed && g.q == 0) a.perm_rows[idx] = a.perm[idx];
    HashFlags f;
    if (in) {
        if (g.q == 0) advance_energies(p, a, idx);
        const uint32_t c = g.q < 3 ? g.q : 0u;
        const double fq = reinterpret_cast<const double*>(a.out + (size_t)idx * 4 + OUT_F)[c];
        const double tq = reinterpret_cast<const double*>(a.out + (size_t)idx * 4 + OUT_T)[c];
        const double xn = integrate_component(p, a, lp, idx, g.q, fq, tq, a.x[idx], a.vel[idx],
                                              a.angvel[idx], a.rank[idx]);
        const double4 x_new = make_double4(g.template bc<0>(xn), g.template bc<1>(xn),
                                           g.template bc<2>(xn), g.template bc<3>(xn));
        if (g.q == 0) hash_particle(p, a, idx, p.L[2], x_new, f);
    }
    publish_next(p, a, lp, f, idx, step_slot(lp) ^ 1u);
}

// The wall-force block partials alone (the reference's summation kernel,
// src/integration.cuh:309-341), after the fused contact-and-integration
// pass: 256 original indices per workgroup.
__global__ __launch_bounds__(BLOCK) void k_wall_partials(const DevArrays a) {
    __shared__ double cache[BLOCK];
    CParams& p = chunk_params(a, blockIdx.x);
    const uint32_t idx = blockIdx.x * BLOCK + threadIdx.x;
    const bool in = idx - p.base < p.np;
    // After the fused pass, whose rows are in the current order.
    if (a.loop[p.sim].order_changed) a.perm_rows[idx] = a.perm[idx];
    const uint32_t nsum = p.nw == 0 ? 0u : (p.wall_sums ? p.nw : 1u);
    for (uint32_t w = 0; w < nsum; ++w) {
        const double v = in ? wall_force_pp(a, idx, w) : 0.0;
        const double s = block_tree_sum(v, cache);
        if (threadIdx.x == 0) a.partial[(size_t)w * a.nchunk + blockIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// The re-sort for the next step inside one workgroup (run at the end of a
// step by k_walls).  When few particles changed cell, the new stable order
// of (cell key, original index) is the old order with those particles taken
// out and merged back at their new keys: the movers are sorted in LDS, each
// one's place among the unchanged particles is a binary search in the old
// order, and every other position copies from the old order at a known
// shift.  Otherwise the workgroup runs the radix passes itself.  Either way
// the result is the same permutation the per-stage sort produces, followed
// by the same cell pass.

template <typename T>
__device__ __forceinline__ void bitonic_sort_lds(T* v, uint32_t n) {
    for (uint32_t k = 2; k <= n; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const T x = v[i], y = v[ixj];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        v[i] = y;
                        v[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
}

// Number of entries of the ascending v[0, n) that are < x.
template <typename T>
__device__ __forceinline__ uint32_t count_below(const T* v, uint32_t n, T x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (v[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ unsigned long long composite(uint32_t key, uint32_t kmask,
                                                        uint32_t idx) {
    return ((unsigned long long)(key & kmask) << 32) | idx;
}

// Re-sort simulation p's segment [B, B + N) of the sorted order (N = npad:
// its particles, then its padding, whose keys never change).
__device__ void resort_one_block(CParams& p, const DevArrays& a, int passes, int bits,
                                 int sorted_bits) {
    __shared__ unsigned long long comp[MV_CAP];  // movers' (new key, index), ascending
    __shared__ uint32_t rk[MV_CAP];              // movers' old positions, ascending
    __shared__ uint32_t npos[MV_CAP];            // movers' new positions
    const uint32_t tid = threadIdx.x;
    const uint32_t B = p.base, N = p.npad;
    const DevLoop* lp = a.loop + p.sim;
    const uint32_t* mv = a.movers + (size_t)p.sim * MV_CAP;
    const uint32_t M = lp->nmovers;
    // The radix sort orders by the low bits * passes bits of the key.
    const uint32_t kmask = bits * passes >= 32 ? 0xffffffffu : ((1u << (bits * passes)) - 1u);
    const bool merge = M <= p.mv_max && M <= MV_CAP;
    // Keep the order and cell ranges the finished step used.
    for (uint32_t j = tid; j < N; j += BLOCK) {
        a.skey_bk[B + j] = a.skey[B + j];
        a.perm_bk[B + j] = a.perm[B + j];
    }
    for (uint32_t c = tid; c < p.ncells; c += BLOCK) {
        a.cs_bk[p.cell_base + c] = a.cell_start[p.cell_base + c];
        a.ce_bk[p.cell_base + c] = a.cell_end[p.cell_base + c];
    }
    __syncthreads();
    if (merge) {
        uint32_t n2 = 1;
        while (n2 < M) n2 <<= 1;
        for (uint32_t m = tid; m < n2; m += BLOCK) {
            if (m < M) {
                const uint32_t o = mv[m];
                comp[m] = composite(a.key[o], kmask, o);
                rk[m] = a.rank[o] - B;
            } else {
                comp[m] = ~0ull;
                rk[m] = 0xffffffffu;
            }
        }
        __syncthreads();
        bitonic_sort_lds(comp, n2);
        bitonic_sort_lds(rk, n2);
        // A mover's new position: the unchanged particles before it (old
        // elements below its composite, minus the movers among them) plus
        // the movers before it.
        for (uint32_t m = tid; m < M; m += BLOCK) {
            const unsigned long long c = comp[m];
            uint32_t lo = 0, hi = N;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (composite(a.skey[B + mid], kmask, a.perm[B + mid]) < c) lo = mid + 1;
                else hi = mid;
            }
            npos[m] = lo - count_below(rk, M, lo) + m;
        }
        __syncthreads();
        // New positions, a contiguous chunk per work-item.  The n-th
        // unchanged particle sits at old position n + d, d = the number of
        // movers with rk[d'] - d' <= n.
        const uint32_t chunk = (N + BLOCK - 1) / BLOCK;
        const uint32_t j0 = tid * chunk;
        const uint32_t j1 = j0 + chunk < N ? j0 + chunk : N;
        if (j0 < j1) {
            uint32_t e = count_below(npos, M, j0);
            uint32_t n = j0 - e;
            uint32_t lo = 0, hi = M;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (rk[mid] - mid <= n) lo = mid + 1; else hi = mid;
            }
            uint32_t d = lo;
            for (uint32_t j = j0; j < j1; ++j) {
                if (e < M && npos[e] == j) {
                    const uint32_t o = (uint32_t)(comp[e] & 0xffffffffull);
                    a.key_tmp[B + j] = a.key[o];
                    a.val_tmp[B + j] = o;
                    ++e;
                } else {
                    while (d < M && rk[d] - d <= n) ++d;
                    const uint32_t i = n + d;
                    a.key_tmp[B + j] = a.skey[B + i];
                    a.val_tmp[B + j] = a.perm[B + i];
                    ++n;
                }
            }
        }
        __syncthreads();
        for (uint32_t j = tid; j < N; j += BLOCK) {
            a.skey[B + j] = a.key_tmp[B + j];
            a.perm[B + j] = a.val_tmp[B + j];
        }
    } else {
        const uint32_t nt = n_tiles(N);
        uint32_t* hist = a.rhist + p.rhist_base;
        for (int pass = 0; pass < passes; ++pass) {
            const RadixPass r = radix_pass(a, pass, passes, B);
            const int shift = bits * pass;
            for (uint32_t t = 0; t < nt; ++t)
                radix_hist_tile(r.kin, N, shift, bits, hist, nt, t);
            __syncthreads();
            radix_scan_block<BLOCK>(hist, (1u << bits) * nt);
            for (uint32_t t = 0; t < nt; ++t)
                radix_scatter_tile(r.kin, r.vin, r.kout, r.vout, N, shift, bits, hist, nt, t, B);
            __syncthreads();
        }
    }
    for (uint32_t c = tid; c < p.ncells; c += BLOCK) a.cell_start[p.cell_base + c] = 0xffffffffu;
    __syncthreads();
    // For the next step (the loop counter advances after this re-sort).  The
    // sorted copies are left in the previous order: gathering them here, in
    // one workgroup, would cost far more than the next step's contact pass
    // reading the original-index arrays through perm (copies_stale), after
    // which the integration pass writes them in the new order.
    const uint32_t nxt = cur_set(a.loop + p.sim) ^ 1u;
    for (uint32_t j = tid; j < N; j += BLOCK) cells_one(p, a, B + j, sorted_bits, nxt, false);
    __syncthreads();
}

// integrateWalls + the loop advance (src/integration.cuh:344-450,
// src/device.cu:2454-2462,2500-2538), one workgroup.  Every dynamic wall
// integrates wall 0's force sum (src/device.cu:2504-2505).  With `resort`
// set it first re-sorts for the next step when the integration pass found
// particles that changed cell (the next step then launches no sort).
__global__ __launch_bounds__(BLOCK) void k_walls(const DevArrays a, int resort, int passes,
                                                 int bits, int sorted_bits) {
    __shared__ double stage[2048];
    __shared__ uint32_t do_sort;
    CParams& p = sim_params(a, blockIdx.x);  // one workgroup per simulation
    DevLoop* lp = a.loop + p.sim;
    if (threadIdx.x == 0) do_sort = (resort && lp->changed[step_slot(lp) ^ 1u]) ? 1u : 0u;
    __shared__ double s_fsum[MAXWALLS];
    const bool walls = p.nw > 0 && p.np > 0;
    // Work-item w integrates wall w (src/integration.cuh:344-450).  Its
    // state, the clock and the normal load are read and formed before the
    // force sums, so that their latency hides behind the sums'.
    const uint32_t w = threadIdx.x;
    const bool mine = walls && w < p.nw && p.wmode[w] != 0;
    const double t_now = lp->t;
    const int64_t iter = lp->iter;
    const double dt = p.dt;
    // The loop state the last work-item updates, read now: after its first
    // store each read would wait for the stores before it.
    const uint32_t slot = (uint32_t)(iter & 1);
    const uint32_t vstate = lp->velocity_state;
    const bool built = p.vl_enabled && (lp->rebuild[slot] || lp->force_rebuild);
    const uint32_t sorts = lp->sorts, builds = lp->vl_builds;
    double4 w_nx = make_double4(0.0, 0.0, 0.0, 0.0), w_mvfd = w_nx;
    double acc0 = 0.0, N = 0.0;
    if (mine) {
        w_nx = a.w_nx[p.wall_base + w];
        w_mvfd = a.w_mvfd[p.wall_base + w];
        if (iter != 0) acc0 = a.w_acc[p.wall_base + w];
        const double sigma_0 = w_mvfd.w + p.sigma0_A * sin(2.0 * PI * p.sigma0_f * t_now);
        N = -sigma_0 * p.L[0] * p.L[1];
    }
    const double grid_top = (p.adaptive == 1 && walls && threadIdx.x == 0)
                                ? a.w_nx[p.wall_base].w : 0.0;
    // The serial sums of the block partials: wall 0's (every wall integrates
    // it, P8), or with per-wall sums each wall's.
    const uint32_t nsum = !walls ? 0u : (p.wall_sums ? p.nw : 1u);
    for (uint32_t k = 0; k < nsum; ++k) {
        const double fs = serial_sum(a.partial + (size_t)k * a.nchunk + p.chunk_base, p.nblocks,
                                     stage);
        if (threadIdx.x == 0) s_fsum[k] = fs;
    }
    __syncthreads();
    if (do_sort) resort_one_block(p, a, passes, bits, sorted_bits);
    if (mine) {
        const uint32_t wi = p.wall_base + w;
        w_mvfd.z = s_fsum[p.wall_sums ? w : 0u];  // 0.0 + partial[0] + partial[1] + ...
        double acc = (w_mvfd.z + N) / w_mvfd.x;
        if (p.wmode[w] == 2) acc = 0.0;
        const double dacc_dt = (acc - acc0) / dt;
        w_nx.w += w_mvfd.y * dt + 0.5 * acc * dt * dt + 1.0 / 6.0 * dacc_dt * dt * dt * dt;
        w_mvfd.y += acc * dt + 0.5 * dacc_dt * dt * dt;
        a.w_nx[wi] = w_nx;
        a.w_mvfd[wi] = w_mvfd;
        a.w_acc[wi] = acc;
    }
    if (threadIdx.x != 0) return;
    if (do_sort) lp->sorts = sorts + 1u;
    lp->resorted_last = do_sort;
    // The next step's contact pass maps its rows to the new order and reads
    // the state through perm.
    lp->order_changed = do_sort;
    lp->copies_stale = do_sort;
    if (p.adaptive == 1 && walls) lp->grid_Lz = grid_top;
    if (t_now >= 10.0 && t_now < 15.0 && vstate == 1) lp->velocity_state = 2;
    else if (t_now >= 15.0 && vstate == 2) lp->velocity_state = 1;
    // Retire this step's flag slot; the next step's slot was filled by the
    // integration pass (or is filled by its own hash launch).
    if (built) lp->vl_builds = builds + 1u;
    lp->changed[slot] = 0u;
    a.glob->changed[slot] = 0u;  // every simulation's block writes the same
    lp->rebuild[slot] = 0u;
    lp->force_rebuild = 0u;
    lp->zero_empty = 0u;
    lp->nmovers = 0u;
    lp->t = t_now + dt;
    lp->iter = iter + 1;
}

// ---------------------------------------------------------------------------
// Bonds.  k_bond_forces evaluates each bond once (src/cohesion.cuh:8-233);
// k_bond_gather adds the contributions to each particle in ascending bond
// order, which is the order the reference's serial reading applies them.

// Bond indices are batch indices (simulation base + the .bin's index).
__global__ __launch_bounds__(BLOCK) void k_bond_forces(const DevArrays a) {
    const uint32_t b = blockIdx.x * BLOCK + threadIdx.x;
    if (b >= a.nbond) return;
    CParams& p = sim_params(a, a.bond_sim[b]);
    const uint32_t bx = a.bonds[2 * b], by = a.bonds[2 * b + 1];
    a.bond_ok[b] = 0u;
    if (a.tape_bond_idx) a.tape_bond_idx[b] = make_uint4(bx, by, 0u, 0u);
    if (!(bx - p.base < p.np)) return;
    const double dt = p.dt;
    const double4 d0 = a.bonds_delta[b], o0 = a.bonds_omega[b];
    if (a.tape_bond_in) {
        // The adjoint's record of the bond's inputs.
        double4* t = a.tape_bond_in + (size_t)8 * b;
        t[0] = a.x[bx];
        t[1] = a.x[by];
        t[2] = a.vel[bx];
        t[3] = a.vel[by];
        t[4] = a.angvel[bx];
        t[5] = a.angvel[by];
        t[6] = d0;
        t[7] = o0;
    }
    d3 delta0_t = mk3(d0.x, d0.y, d0.z);
    const double delta0_n = d0.w;
    d3 omega0_t = mk3(o0.x, o0.y, o0.z);
    const double omega0_n = o0.w;
    const double4 x_i = a.x[bx], x_j = a.x[by];
    const double4 vel_i = a.vel[bx], vel_j = a.vel[by];
    const d3 angvel_i = xyz(a.angvel[bx]), angvel_j = xyz(a.angvel[by]);
    const double R_bar = p.lambda_bar * fmin(x_i.w, x_j.w);
    const double A = PI * R_bar * R_bar;
    const double I = 0.25 * PI * R_bar * R_bar * R_bar * R_bar;
    const double J = I * 2.0;
    const d3 x = mk3(x_i.x - x_j.x, x_i.y - x_j.y, x_i.z - x_j.z);
    const double x_length = length(x);
    const double overlap = fmin(0.0, x_length - (x_i.w + x_j.w));
    const d3 n = x / x_length;
    delta0_t = delta0_t - (n * dot(n, delta0_t));
    const d3 ddelta = (mk3(vel_i.x - vel_j.x, vel_i.y - vel_j.y, vel_i.z - vel_j.z) +
                       (x_i.w + overlap / 2.0) * cross(n, angvel_i) +
                       (x_j.w + overlap / 2.0) * cross(n, angvel_j)) * dt;
    const double ddelta_n = -dot(ddelta, n);
    const double delta_n = delta0_n + ddelta_n;
    const d3 ddelta_t = ddelta - n * dot(n, ddelta);
    const d3 delta_t = delta0_t + ddelta_t;
    // Stiffness per unit area: the contact's k_n, k_t, or the bond option.
    const double kn = p.bond_k > 0.0 ? p.bond_k / A : p.k_n;
    const double kt = p.bond_k > 0.0 ? p.bond_k / A : p.k_t;
    const d3 f_n = (kn * A * delta_n + p.gamma_n * ddelta_n / dt) * n;
    const d3 f_t = -kt * A * delta_t - p.gamma_t * ddelta_t / dt;
    const d3 f = f_n + f_t;
    const d3 mn = -n;
    omega0_t = omega0_t - mn * dot(omega0_t, mn);
    const d3 domega = mk3(angvel_j.x - angvel_i.x, angvel_j.y - angvel_i.y,
                          angvel_j.z - angvel_i.z) * dt;
    const double domega_n = dot(domega, mn);
    const double omega_n = omega0_n + domega_n;
    const d3 domega_t = domega - domega_n * mn;
    const d3 omega_t = omega0_t + domega_t;
    const d3 t_n = -kt * J * omega_n * mn;
    const d3 t_t = -kn * I * omega_t;
    const d3 t_j = t_n + t_t;
    bool breaks;
    if (p.bond_break_ext > 0.0) {
        // The option: a normal extension (delta_n < 0 when the pair
        // separates) of bond_break_ext times the smaller radius.
        breaks = -delta_n >= p.bond_break_ext * fmin(x_i.w, x_j.w);
    } else {
        const double sigma_max = length(f_n) / A + length(t_t) * R_bar / I;
        const double tau_max = length(f_t) / A + length(t_n) * R_bar / J;
        breaks = sigma_max >= p.sigma_b || tau_max >= p.tau_b;
    }
    if (breaks) {
        a.bonds[2 * b] = p.base + p.nb0;  // src/cohesion.cuh:205
        if (p.nb0 < p.np) atomicOr(&a.loop[p.sim].bond_retargeted, 1u);
        return;
    }
    a.bonds_delta[b] = make_double4(delta_t.x, delta_t.y, delta_t.z, delta_n);
    a.bonds_omega[b] = make_double4(omega_t.x, omega_t.y, omega_t.z, omega_n);
    a.bond_f[b] = make_double4(f.x, f.y, f.z, 0.0);
    a.bond_t[b] = make_double4(t_j.x, t_j.y, t_j.z, 0.0);
    a.bond_ok[b] = 1u;
    if (a.tape_bond_idx) a.tape_bond_idx[b] = make_uint4(bx, by, 1u, 0u);
}

// Per-particle gather over the bonds that touch the particle, in ascending
// bond id (CSR of the initial bond list, built at open).  A broken bond has
// its first index reset to nb0 (src/cohesion.cuh:205); once that has
// happened, particle nb0 scans the whole bond list instead.
__device__ __forceinline__ void apply_bond(const DevArrays& a, uint32_t i, uint32_t b,
                                           double4& F, double4& T, bool& touched) {
    if (!a.bond_ok[b]) return;
    const uint32_t bx = a.bonds[2 * b], by = a.bonds[2 * b + 1];
    if (bx != i && by != i) return;
    const double4 f = a.bond_f[b], t = a.bond_t[b];
    // xyz only: the w parts of the record hold es and ev.
    if (bx == i) {
        F.x += f.x; F.y += f.y; F.z += f.z;
        T.x -= t.x; T.y -= t.y; T.z -= t.z;
    }
    if (by == i) {
        F.x -= f.x; F.y -= f.y; F.z -= f.z;
        T.x += t.x; T.y += t.y; T.z += t.z;
    }
    touched = true;
}

__global__ __launch_bounds__(BLOCK) void k_bond_gather(const DevArrays a) {
    CParams& p = chunk_params(a, blockIdx.x);
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i - p.base >= p.np) return;
    const bool full = (i - p.base == p.nb0) && a.loop[p.sim].bond_retargeted != 0u;
    if (a.tape_bond_full && i - p.base == p.nb0) a.tape_bond_full[p.sim] = full ? 1u : 0u;
    const uint32_t b0 = full ? p.bond_base : a.badj_off[i];
    const uint32_t b1 = full ? p.bond_base + p.nb0 : a.badj_off[i + 1];
    if (b0 == b1) return;
    // Force and torque of the output record (their w parts, es and ev,
    // pass through unchanged).
    double4* const rec = a.out + (size_t)i * 4;
    double4 F = rec[OUT_F], T = rec[OUT_T];
    bool touched = false;
    for (uint32_t k = b0; k < b1; ++k) apply_bond(a, i, full ? k : a.badj[k], F, T, touched);
    if (touched) {
        rec[OUT_F] = F;
        rec[OUT_T] = T;
    }
}

// Self-check of sqrt_rn_f32: the FP32 values with bit patterns in [lo, hi)
// against the reference root; mismatches (other than two NaNs) are counted.
__global__ __launch_bounds__(BLOCK) void k_check_sqrt32(unsigned long long* bad, uint64_t lo,
                                                        uint64_t hi) {
    unsigned long long n = 0;
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (uint64_t u = lo + (uint64_t)blockIdx.x * BLOCK + threadIdx.x; u < hi; u += stride) {
        const float x = __uint_as_float((uint32_t)u);
        const float y = sqrt_rn_f32(x), r = sqrt_rn_f32_ref(x);
        const bool nan2 = y != y && r != r;
        n += (!nan2 && __float_as_uint(y) != __float_as_uint(r)) ? 1ull : 0ull;
    }
    if (n) atomicAdd(bad, n);
}

}  // namespace

// Contact rows whose words index the order of perm_rows, rewritten with the
// partners' original indices.
__global__ __launch_bounds__(BLOCK) void k_rows_to_orig(const uint32_t* __restrict__ rows,
                                                       const uint32_t* __restrict__ perm_rows,
                                                       uint32_t* __restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t w = rows[i], idx = w & CW_IDX_MASK;
    out[i] = idx == EMPTY ? w : (w & ~CW_IDX_MASK) | perm_rows[idx];
}

// ---------------------------------------------------------------------------
// Launch wrappers

hipError_t launch_fill_u32(uint32_t* dst, unsigned n, uint32_t v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(n_blocks(n)), dim3(BLOCK), 0, s, nullptr, nullptr, dst,
                       n, v);
    return hipGetLastError();
}

hipError_t launch_hash(const DevArrays& a, hipStream_t s) {
    hipLaunchKernelGGL(k_hash, dim3(a.nchunk), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_sort(const DevArrays& a, int passes, int bits, hipStream_t s) {
    const uint32_t n = a.ntot;
    const uint32_t nt = n_tiles(n);
    for (int pass = 0; pass < passes; ++pass) {
        const RadixPass r = radix_pass(a, pass, passes, 0u);
        const int shift = bits * pass;
        hipLaunchKernelGGL(k_radix_hist, dim3(nt), dim3(BLOCK), 0, s, a.glob, a.loop, r.kin, n,
                           shift, bits, a.rhist, nt);
        hipLaunchKernelGGL(k_radix_scan, dim3(1), dim3(1024), 0, s, a.glob, a.loop, a.rhist,
                           (1u << bits) * nt);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nt), dim3(BLOCK), 0, s, a.glob, a.loop, r.kin,
                           r.vin, r.kout, r.vout, n, shift, bits, a.rhist, nt);
    }
    return hipGetLastError();
}

hipError_t launch_cells(const DevArrays& a, int sorted_bits, hipStream_t s) {
    if (a.nkeys)
        hipLaunchKernelGGL(k_fill_u32, dim3(n_blocks(a.nkeys)), dim3(BLOCK), 0, s, a.glob, a.loop,
                           a.cell_start, a.nkeys, 0xffffffffu);
    hipLaunchKernelGGL(k_cells, dim3(a.nchunk), dim3(BLOCK), 0, s, a, sorted_bits);
    return hipGetLastError();
}

template <int CM>
void launch_contact_lanes(const DevArrays& a, int lanes, int wg, hipStream_t s) {
    // BLOCK / (wg / lanes) workgroups per 256-particle chunk.
    if (lanes == 4) {
        const unsigned per = (unsigned)(BLOCK * 4 / wg);
        if (wg == 128)
            hipLaunchKernelGGL((k_contact_g<CM, 4, false, 128>), dim3(a.nchunk * per), dim3(128),
                               0, s, a);
        else if (wg == 512)
            hipLaunchKernelGGL((k_contact_g<CM, 4, false, 512>), dim3(a.nchunk * per), dim3(512),
                               0, s, a);
        else
            hipLaunchKernelGGL((k_contact_g<CM, 4, false, 256>), dim3(a.nchunk * per), dim3(256),
                               0, s, a);
    } else if (lanes == 8) {
        hipLaunchKernelGGL((k_contact_g<CM, 8, false, BLOCK>), dim3(a.nchunk * 8u), dim3(BLOCK), 0,
                           s, a);
    } else {
        hipLaunchKernelGGL((k_contact_g<CM, 16, false, BLOCK>), dim3(a.nchunk * 16u), dim3(BLOCK),
                           0, s, a);
    }
}

hipError_t launch_contact(const DevArrays& a, int contactmodel, int lanes, int wg,
                          hipStream_t s) {
    const dim3 b(BLOCK);
    if (lanes > 1 && (contactmodel == 2 || contactmodel == 3)) {
        if (contactmodel == 2) launch_contact_lanes<2>(a, lanes, wg, s);
        else launch_contact_lanes<3>(a, lanes, wg, s);
        return hipGetLastError();
    }
    const dim3 g(a.nchunk);
    if (contactmodel == 1)
        hipLaunchKernelGGL(k_contact<1>, g, b, 0, s, a);
    else if (contactmodel == 3)
        hipLaunchKernelGGL(k_contact<3>, g, b, 0, s, a);
    else if (contactmodel == 2)
        hipLaunchKernelGGL(k_contact<2>, g, b, 0, s, a);
    else
        hipLaunchKernelGGL(k_contact<0>, g, b, 0, s, a);
    return hipGetLastError();
}

hipError_t launch_contact_integrate(const DevArrays& a, int contactmodel, hipStream_t s) {
    const dim3 b(BLOCK), g(a.nchunk * 4u);
    if (contactmodel == 2)
        hipLaunchKernelGGL((k_contact_g<2, 4, true, BLOCK>), g, b, 0, s, a);
    else
        hipLaunchKernelGGL((k_contact_g<3, 4, true, BLOCK>), g, b, 0, s, a);
    return hipGetLastError();
}

hipError_t launch_wall_partials(const DevArrays& a, hipStream_t s) {
    hipLaunchKernelGGL(k_wall_partials, dim3(a.nchunk), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bonds(const DevArrays& a, hipStream_t s) {
    if (a.nbond == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bond_forces, dim3(n_blocks(a.nbond)), dim3(BLOCK), 0, s, a);
    hipLaunchKernelGGL(k_bond_gather, dim3(a.nchunk), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_shear_stress(const DevArrays& a, hipStream_t s) {
    hipLaunchKernelGGL(k_shear_stress, dim3(a.nchunk), dim3(BLOCK), 0, s, a);
    hipLaunchKernelGGL(k_tau_sum, dim3(a.nsim), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_integrate(const DevArrays& a, int hash_next, hipStream_t s) {
    hipLaunchKernelGGL(k_integrate, dim3(a.nchunk), dim3(BLOCK), 0, s, a, hash_next);
    return hipGetLastError();
}

hipError_t launch_integrate_q(const DevArrays& a, hipStream_t s) {
    hipLaunchKernelGGL(k_integrate_q, dim3(a.nchunk * 4), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_walls(const DevArrays& a, int resort, int passes, int bits, int sorted_bits,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_walls, dim3(a.nsim), dim3(BLOCK), 0, s, a, resort, passes, bits,
                       sorted_bits);
    return hipGetLastError();
}

hipError_t launch_rows_to_orig(const uint32_t* rows, const uint32_t* perm_rows, uint32_t* out,
                               uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rows_to_orig, dim3(n_blocks(n)), dim3(BLOCK), 0, s, rows, perm_rows,
                       out, n);
    return hipGetLastError();
}

hipError_t launch_check_sqrt32(unsigned long long* bad, uint64_t lo, uint64_t hi, hipStream_t s) {
    hipLaunchKernelGGL(k_check_sqrt32, dim3(4096), dim3(BLOCK), 0, s, bad, lo, hi);
    return hipGetLastError();
}

}  // namespace sphx
