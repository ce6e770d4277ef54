// Host side of the sphere_hip C ABI: context lifetime, .bin I/O, input
// validation, the step loop and the output loop of DEM::startTime.
//
// Reference behaviour followed here: DEM::DEM and checkValues
// (src/sphere.cpp:16-95,145-345), readbin/writebin (src/file_io.cpp:30-731),
// DEM::startTime (src/device.cu:826-2700).  The library reports errors
// through return codes and sph_last_error() instead of exit().

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "../../include/sphere_hip.h"
#include "dem_arrays.hpp"
#include "adjoint.hpp"
#include "analysis.hpp"
#include "dem_kernels.hpp"
#include "dem_types.hpp"

#include <mutex>

// The staging array of sph_porosity_noalloc: z positions then porosities of
// up to SPH_POROSITY_NOALLOC_MAX_SLICES slices, one copy on every GPU the
// library runs on, held until the process ends.  Host code uses only the
// device address that hipGetSymbolAddress returns for it (per device).
static __device__ __align__(256) double scratch[2 * SPH_POROSITY_NOALLOC_MAX_SLICES];

using namespace sphx;

namespace {

thread_local std::string g_open_error;

struct Host {
    // header (src/file_io.cpp:47-114)
    double version = SPH_VERSION;
    int32_t nd = 3;
    uint32_t np = 0;
    double dt = 0, t = 0, t_total = 0, file_dt = 0;
    uint32_t step_count = 0;
    double origo[3] = {0, 0, 0}, L[3] = {0, 0, 0};
    uint32_t num[3] = {0, 0, 0};
    int32_t periodic = 0, adaptive = 0;
    // particles, AoS 4 doubles as on the device
    std::vector<double> x, xyzsum, vel, force, angpos, angvel, torque;
    std::vector<double> es_dot, es, ev_dot, ev, p;
    std::vector<int32_t> color;
    // params (src/datatypes.h:74-101)
    double g[3] = {0, 0, 0};
    double k_n = 0, k_t = 0, k_r = 0, E = 0, gamma_n = 0, gamma_t = 0, gamma_r = 0;
    double mu_s = 0, mu_d = 0, mu_r = 0, gamma_wn = 0, gamma_wt = 0, mu_ws = 0, mu_wd = 0;
    double rho = 0;
    uint32_t contactmodel = 0;
    double kappa = 0, db = 0, V_b = 0;
    uint32_t nw = 0;
    int32_t wmode[MAXWALLS] = {0, 0, 0, 0, 0, 0};
    std::vector<double> w_nx, w_mvfd;  // nw x 4
    double sigma0_A = 0, sigma0_f = 0, tau_x = 0;
    double lambda_bar = 1, sigma_b = 0, tau_b = 0;
    uint32_t nb0 = 0;
    std::vector<uint32_t> bonds;       // nb0 x 2
    std::vector<double> bonds_delta, bonds_omega;  // nb0 x 4 (t.x, t.y, t.z, n)
};

struct Reader {
    const unsigned char* p;
    size_t n, off;
    bool ok;
    template <class T> T get() {
        T v;
        if (off + sizeof(T) > n) {
            ok = false;
            memset(&v, 0, sizeof(T));
            return v;
        }
        memcpy(&v, p + off, sizeof(T));
        off += sizeof(T);
        return v;
    }
    void getv(double* dst, size_t k) { for (size_t i = 0; i < k; ++i) dst[i] = get<double>(); }
};

struct Writer {
    std::vector<unsigned char> b;
    template <class T> void put(const T& v) {
        const unsigned char* q = reinterpret_cast<const unsigned char*>(&v);
        b.insert(b.end(), q, q + sizeof(T));
    }
};

}  // namespace

struct sph_prof_state;

// One simulation of a context: its .bin state on the host, its parameter
// block (also in the device table), and the host mirror of its loop.
struct Sim {
    Host h;
    DevParams P;
    int64_t iter = 0;
};

// The adjoint's host bookkeeping at a checkpoint: the loop mirrors and the
// hash/re-sort state enqueue_step reads.
struct AdjHostMark {
    std::vector<double> t;
    std::vector<int64_t> iter;
    bool hash_pending, prev_walls_resorted;
};

// A recorded forward pass awaiting its reverse (sph_adjoint_record).
struct AdjPlan {
    uint64_t nsteps = 0, every = 0, K = 0;
    int64_t iter0 = 0, iter_end = 0;
    void* buf = nullptr;
    std::vector<AdjHostMark> marks;
};

struct sph_ctx {
    std::vector<Sim> sims;  // one for sph_open, several for sph_open_batch
    Host& h0() { return sims[0].h; }
    DevArrays A;
    DevParams* d_ptab = nullptr;  // device copy of the parameter blocks
    std::vector<void*> allocs;
    std::vector<size_t> alloc_bytes;  // size of each allocation, in order
    hipStream_t stream = nullptr;
    int device = 0;
    unsigned flags = 0;
    std::string err;
    // Sort width over the batch's global keys.
    int sorted_bits = 32;
    int passes = 4;
    uint32_t contactmodel = 0;   // shared by all simulations of a batch
    bool shear_stress = false;   // some simulation has a wmode-3 top wall
    bool any_adaptive = false;   // some simulation has an adaptive grid
    bool profile = false;
    // The next step starts with a hash launch (keys not yet computed from
    // the current positions); otherwise the previous integration pass left
    // them, except with an adaptive grid, which hashes every step.
    bool hash_pending = true;
    // Work-items per particle in the contact kernel (models 2 and 3).
    int contact_lanes = 4;
    int contact_wg = 256;     // SPH_S_CONTACT_WG (4 lanes)
    int radix_bits = 8;       // SPH_S_RADIX_BITS
    // Where a step's re-sort runs: 0 = the per-stage sort launches at the
    // step start (every step; they return at once when nothing changed),
    // 1 = one workgroup per simulation at the end of the previous step, only
    // when needed, 2 = choose between the two from the measured rate.
    int sort_mode = 2;
    bool resort_at_end = true;     // the choice for the steps being enqueued
    bool prev_walls_resorted = false;
    // Asynchronous sample of the loop counters for the automatic choice.
    DevLoop* loop_sample = nullptr;   // pinned host memory (all simulations)
    hipEvent_t sample_ev = nullptr;
    bool sample_pending = false;
    int64_t sample_iter0 = 0;
    uint64_t sample_sorts0 = 0;
    // Steady-state step chunks replayed from captured HIP graphs (0 = off):
    // one graph per (chunk length, re-sort placement, contact lanes).
    int graph_steps = 0;
    struct Graph {
        int steps;
        bool resort;
        int lanes;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    // While an adjoint forward pass records: where the contact rows after
    // the interaction pass go.
    // The adjoint's tape and reverse state, kept between calls, and its
    // segment length (0: automatic).
    void* adj_buf = nullptr;
    size_t adj_bytes = 0;
    AdjPlan* adj_plan = nullptr;  // a recorded forward pass awaiting its reverse
    uint64_t adj_checkpoint = 0;
    // Contact and integration in one pass where the step allows it (off by
    // default: measured slower than the two kernels, 72.1 + 6.2 us for the
    // fused pass and its wall partials against 53.3 + 19.6 us).
    bool fused = false;
    int integrate_lanes = 1;  // SPH_S_INTEGRATE_LANES
    // per-kernel profile (stage name, total ms, launches)
    std::vector<std::string> prof_names;
    std::vector<double> prof_ms;
    std::vector<long> prof_n;
    sph_prof_state* prof = nullptr;
};

namespace {

int fail(sph_ctx* c, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else g_open_error = buf;
    return code;
}

int hip_fail(sph_ctx* c, hipError_t e, const char* where) {
    return fail(c, SPH_E_DEVICE, "%s: %s", where, hipGetErrorString(e));
}

#define HIPCHK(c, expr, where)                        \
    do {                                              \
        hipError_t e_ = (expr);                       \
        if (e_ != hipSuccess) return hip_fail(c, e_, where); \
    } while (0)

// .bin v2.15 reader, fluid == 0 (src/file_io.cpp:30-418).
int parse_bin(Host& h, const void* bytes, size_t nbytes, std::string& err) {
    Reader r = {static_cast<const unsigned char*>(bytes), nbytes, 0, true};
    h.version = r.get<double>();
    if (!r.ok) { err = "Could not read input binary file (empty)"; return SPH_E_IO; }
    if (h.version != SPH_VERSION) {
        char m[256];
        snprintf(m, sizeof m,
                 "Error: The input file is written by sphere version %g, which is "
                 "incompatible with this version (%g).", h.version, SPH_VERSION);
        err = m;
        return SPH_E_VERSION;
    }
    h.nd = r.get<int32_t>();
    h.np = r.get<uint32_t>();
    if (h.nd != 3) {
        err = "Dimensionality mismatch between dataset and this SPHERE program.";
        return SPH_E_INPUT;
    }
    h.dt = r.get<double>(); h.t = r.get<double>(); h.t_total = r.get<double>();
    h.file_dt = r.get<double>(); h.step_count = r.get<uint32_t>();
    r.getv(h.origo, 3); r.getv(h.L, 3);
    for (int d = 0; d < 3; ++d) h.num[d] = r.get<uint32_t>();
    h.periodic = r.get<int32_t>(); h.adaptive = r.get<int32_t>();
    const size_t np = h.np;
    if (!r.ok || np * 228 > nbytes - std::min(nbytes, r.off)) {
        err = "short read: file too small for np particles";
        return SPH_E_IO;
    }
    h.x.assign(np * 4, 0.0); h.xyzsum.assign(np * 4, 0.0); h.vel.assign(np * 4, 0.0);
    h.force.assign(np * 4, 0.0); h.angpos.assign(np * 4, 0.0); h.angvel.assign(np * 4, 0.0);
    h.torque.assign(np * 4, 0.0);
    for (size_t i = 0; i < np; ++i) r.getv(&h.x[4 * i], 4);
    for (size_t i = 0; i < np; ++i) r.getv(&h.xyzsum[4 * i], 3);
    for (size_t i = 0; i < np; ++i) r.getv(&h.vel[4 * i], 4);
    for (size_t i = 0; i < np; ++i) r.getv(&h.force[4 * i], 3);
    for (size_t i = 0; i < np; ++i) r.getv(&h.angpos[4 * i], 3);
    for (size_t i = 0; i < np; ++i) r.getv(&h.angvel[4 * i], 3);
    for (size_t i = 0; i < np; ++i) r.getv(&h.torque[4 * i], 3);
    std::vector<double>* e[5] = {&h.es_dot, &h.es, &h.ev_dot, &h.ev, &h.p};
    for (int k = 0; k < 5; ++k) { e[k]->assign(np, 0.0); r.getv(e[k]->data(), np); }
    r.getv(h.g, 3);
    h.k_n = r.get<double>(); h.k_t = r.get<double>(); h.k_r = r.get<double>();
    h.E = r.get<double>(); h.gamma_n = r.get<double>(); h.gamma_t = r.get<double>();
    h.gamma_r = r.get<double>(); h.mu_s = r.get<double>(); h.mu_d = r.get<double>();
    h.mu_r = r.get<double>(); h.gamma_wn = r.get<double>(); h.gamma_wt = r.get<double>();
    h.mu_ws = r.get<double>(); h.mu_wd = r.get<double>(); h.rho = r.get<double>();
    h.contactmodel = r.get<uint32_t>();
    h.kappa = r.get<double>(); h.db = r.get<double>(); h.V_b = r.get<double>();
    h.nw = r.get<uint32_t>();
    if (!r.ok) { err = "short read"; return SPH_E_IO; }
    if (h.nw > (uint32_t)MAXWALLS) {
        char m[160];
        snprintf(m, sizeof m, "Error; MAXWALLS (%d) in datatypes.h is smaller than the number "
                 "of walls specified in the input file (%u).", MAXWALLS, h.nw);
        err = m;
        return SPH_E_INPUT;
    }
    for (uint32_t i = 0; i < h.nw; ++i) h.wmode[i] = r.get<int32_t>();
    h.w_nx.assign(h.nw * 4, 0.0); h.w_mvfd.assign(h.nw * 4, 0.0);
    for (uint32_t i = 0; i < h.nw; ++i) r.getv(&h.w_nx[4 * i], 4);
    for (uint32_t i = 0; i < h.nw; ++i) r.getv(&h.w_mvfd[4 * i], 4);
    h.sigma0_A = r.get<double>(); h.sigma0_f = r.get<double>(); h.tau_x = r.get<double>();
    h.lambda_bar = r.get<double>(); h.nb0 = r.get<uint32_t>();
    h.sigma_b = r.get<double>(); h.tau_b = r.get<double>();
    if (!r.ok || (size_t)h.nb0 * 72 > nbytes - std::min(nbytes, r.off)) {
        err = "short read";
        return SPH_E_IO;
    }
    const size_t nb = h.nb0;
    h.bonds.assign(nb * 2, 0u); h.bonds_delta.assign(nb * 4, 0.0); h.bonds_omega.assign(nb * 4, 0.0);
    for (size_t i = 0; i < nb; ++i) { h.bonds[2 * i] = r.get<uint32_t>(); h.bonds[2 * i + 1] = r.get<uint32_t>(); }
    for (size_t i = 0; i < nb; ++i) h.bonds_delta[4 * i + 3] = r.get<double>();
    for (size_t i = 0; i < nb; ++i) r.getv(&h.bonds_delta[4 * i], 3);
    for (size_t i = 0; i < nb; ++i) h.bonds_omega[4 * i + 3] = r.get<double>();
    for (size_t i = 0; i < nb; ++i) r.getv(&h.bonds_omega[4 * i], 3);
    h.color.assign(np, 0);
    for (size_t i = 0; i < np; ++i) h.color[i] = r.get<int32_t>();
    if (!r.ok) { err = "short read"; return SPH_E_IO; }
    return SPH_OK;
}

void write_bin(const Host& h, Writer& w) {
    w.put(h.version); w.put(h.nd); w.put(h.np);
    w.put(h.dt); w.put(h.t); w.put(h.t_total); w.put(h.file_dt); w.put(h.step_count);
    for (int d = 0; d < 3; ++d) w.put(h.origo[d]);
    for (int d = 0; d < 3; ++d) w.put(h.L[d]);
    for (int d = 0; d < 3; ++d) w.put(h.num[d]);
    w.put(h.periodic); w.put(h.adaptive);
    const size_t np = h.np;
    w.b.reserve(w.b.size() + np * 228 + 4096);
    for (size_t i = 0; i < np; ++i) for (int k = 0; k < 4; ++k) w.put(h.x[4 * i + k]);
    for (size_t i = 0; i < np; ++i) for (int k = 0; k < 3; ++k) w.put(h.xyzsum[4 * i + k]);
    for (size_t i = 0; i < np; ++i) for (int k = 0; k < 4; ++k) w.put(h.vel[4 * i + k]);
    const std::vector<double>* v3[4] = {&h.force, &h.angpos, &h.angvel, &h.torque};
    for (int a = 0; a < 4; ++a)
        for (size_t i = 0; i < np; ++i) for (int k = 0; k < 3; ++k) w.put((*v3[a])[4 * i + k]);
    const std::vector<double>* e[5] = {&h.es_dot, &h.es, &h.ev_dot, &h.ev, &h.p};
    for (int a = 0; a < 5; ++a) for (size_t i = 0; i < np; ++i) w.put((*e[a])[i]);
    for (int d = 0; d < 3; ++d) w.put(h.g[d]);
    w.put(h.k_n); w.put(h.k_t); w.put(h.k_r); w.put(h.E); w.put(h.gamma_n); w.put(h.gamma_t);
    w.put(h.gamma_r); w.put(h.mu_s); w.put(h.mu_d); w.put(h.mu_r); w.put(h.gamma_wn);
    w.put(h.gamma_wt); w.put(h.mu_ws); w.put(h.mu_wd); w.put(h.rho); w.put(h.contactmodel);
    w.put(h.kappa); w.put(h.db); w.put(h.V_b);
    w.put(h.nw);
    for (uint32_t i = 0; i < h.nw; ++i) w.put(h.wmode[i]);
    for (uint32_t i = 0; i < h.nw; ++i) for (int k = 0; k < 4; ++k) w.put(h.w_nx[4 * i + k]);
    for (uint32_t i = 0; i < h.nw; ++i) for (int k = 0; k < 4; ++k) w.put(h.w_mvfd[4 * i + k]);
    w.put(h.sigma0_A); w.put(h.sigma0_f); w.put(h.tau_x);
    w.put(h.lambda_bar); w.put(h.nb0); w.put(h.sigma_b); w.put(h.tau_b);
    const size_t nb = h.nb0;
    for (size_t i = 0; i < nb; ++i) { w.put(h.bonds[2 * i]); w.put(h.bonds[2 * i + 1]); }
    for (size_t i = 0; i < nb; ++i) w.put(h.bonds_delta[4 * i + 3]);
    for (size_t i = 0; i < nb; ++i) for (int k = 0; k < 3; ++k) w.put(h.bonds_delta[4 * i + k]);
    for (size_t i = 0; i < nb; ++i) w.put(h.bonds_omega[4 * i + 3]);
    for (size_t i = 0; i < nb; ++i) for (int k = 0; k < 3; ++k) w.put(h.bonds_omega[4 * i + k]);
    for (size_t i = 0; i < np; ++i) w.put(h.color[i]);
}

// checkValues, src/sphere.cpp:145-345 (fluid == 0).
int check_values(sph_ctx* c, const Host& h) {
    if (h.t > h.t_total || h.t < 0.0)
        return fail(c, SPH_E_INPUT, "Error: time.current = %g s, time.total = %g s", h.t, h.t_total);
    if (h.origo[0] < 0.0 || h.origo[1] < 0.0 || h.origo[2] < 0.0)
        return fail(c, SPH_E_INPUT,
                    "Error: Negative values in grid.origo are known to cause problems. "
                    "grid.origo[0] = %g m, grid.origo[1] = %g m, grid.origo[2] = %g m.",
                    h.origo[0], h.origo[1], h.origo[2]);
    if (h.L[0] <= 0.0 || h.L[1] <= 0.0 || h.L[2] <= 0.0)
        return fail(c, SPH_E_INPUT, "Error: grid.L[0] = %g m, grid.L[1] = %g m, grid.L[2] = %g m.",
                    h.L[0], h.L[1], h.L[2]);
    if (h.num[0] == 0 || h.num[1] == 0 || h.num[2] == 0)
        return fail(c, SPH_E_INPUT, "Error: grid.num[0] = %u, grid.num[1] = %u, grid.num[2] = %u.",
                    h.num[0], h.num[1], h.num[2]);
    if (h.periodic == 2 && h.num[0] < 3)
        return fail(c, SPH_E_INPUT, "Error: When 1st dimension boundaries are periodic, there "
                    "must be at least 3 cells in that dimension.");
    if (h.periodic == 1 && (h.num[0] < 3 || h.num[1] < 3))
        return fail(c, SPH_E_INPUT, "Error: When 1st and 2nd dimension boundaries are periodic, "
                    "there must be at least 3 cells in each of those dimensions.");
    for (uint32_t i = 0; i < h.np; ++i) {
        const double* x = &h.x[4 * i];
        const double* v = &h.vel[4 * i];
        if (x[3] <= 0.0)
            return fail(c, SPH_E_INPUT, "Error: Particle %u has a radius of %g m.", i, x[3]);
        for (int k = 0; k < 4; ++k)
            if (x[k] != x[k])
                return fail(c, SPH_E_INPUT, "Error: Particle %u has a non-real position", i);
        for (int k = 0; k < 4; ++k)
            if (v[k] != v[k])
                return fail(c, SPH_E_INPUT, "Error: Particle %u has a non-real velocity", i);
        if (x[0] < h.origo[0] || x[1] < h.origo[1] || x[2] < h.origo[2] || x[0] > h.L[0] ||
            x[1] > h.L[1] || x[2] > h.L[2])
            return fail(c, SPH_E_INPUT,
                        "Error: Particle %u is outside of the computational grid\n"
                        "k.x[i] = [%g, %g, %g]\ngrid.origo = [%g, %g, %g], grid.L = [%g, %g, %g].",
                        i, x[0], x[1], x[2], h.origo[0], h.origo[1], h.origo[2], h.L[0], h.L[1], h.L[2]);
    }
    if (h.nw > 0) {
        double z_max = 0.0;
        for (uint32_t i = 0; i < h.np; ++i)
            if (h.x[4 * i + 2] > z_max) z_max = h.x[4 * i + 2];
        if (h.w_nx[3] < z_max)
            return fail(c, SPH_E_INPUT,
                        "Error: One or more particles have centres above the upper, dynamic wall");
    }
    if (h.k_n <= 0.0) return fail(c, SPH_E_INPUT, "Error: k_n = %g N/m", h.k_n);
    if (h.E < 0.0) return fail(c, SPH_E_INPUT, "Error: E = %g N/m^3", h.E);
    if (h.rho <= 0.0) return fail(c, SPH_E_INPUT, "Error: rho = %g kg/m3", h.rho);
    return SPH_OK;
}

void set_skin(DevParams& P, double skin) {
    P.vl_skin = skin > 0.0 ? skin : 0.0;
    P.vl_enabled = skin > 0.0 ? 1 : 0;
    // Rebuild once any particle has moved 0.45 skin (< skin / 2) since the
    // lists were built.
    const double trig = 0.45 * P.vl_skin;
    P.vl_trigger2 = trig * trig;
}

int bit_length(uint64_t v) {
    int b = 0;
    while (v) { ++b; v >>= 1; }
    return b;
}

// The physical and grid parameters of one simulation (its place in the
// batch is set by layout()).
void fill_params(Sim& sim) {
    const Host& h = sim.h;
    DevParams& P = sim.P;
    memset(&P, 0, sizeof P);
    for (int d = 0; d < 3; ++d) {
        P.origo[d] = h.origo[d];
        P.L[d] = h.L[d];
        P.num[d] = h.num[d];
        P.g[d] = h.g[d];
    }
    P.periodic = h.periodic;
    P.adaptive = h.adaptive;
    P.ncells = h.num[0] * h.num[1] * h.num[2];
    P.np = h.np;
    P.nw = h.nw;
    P.nb0 = h.nb0;
    P.nblocks = (h.np + BLOCK - 1) / BLOCK;
    P.dt = h.dt;
    P.k_n = h.k_n; P.k_t = h.k_t; P.k_r = h.k_r; P.E = h.E;
    P.gamma_n = h.gamma_n; P.gamma_t = h.gamma_t; P.gamma_r = h.gamma_r;
    P.mu_s = h.mu_s; P.mu_d = h.mu_d; P.mu_r = h.mu_r;
    P.gamma_wn = h.gamma_wn; P.gamma_wt = h.gamma_wt; P.mu_ws = h.mu_ws; P.mu_wd = h.mu_wd;
    P.rho = h.rho;
    P.contactmodel = h.contactmodel;
    P.kappa = h.kappa; P.db = h.db; P.V_b = h.V_b;
    P.lambda_bar = h.lambda_bar; P.sigma_b = h.sigma_b; P.tau_b = h.tau_b;
    P.sigma0_A = h.sigma0_A; P.sigma0_f = h.sigma0_f;
    P.tau_x = h.tau_x;
    for (int w = 0; w < MAXWALLS; ++w) P.wmode[w] = h.wmode[w];
    P.mv_max = MV_CAP;
    // Candidate lists: skin of 0.1 r_min by default.  Model 1 with a
    // capillary debonding distance keeps the full scan (its interaction
    // range extends beyond overlap).
    double rmin = 0.0;
    for (uint32_t i = 0; i < h.np; ++i)
        if (i == 0 || h.x[4 * i + 3] < rmin) rmin = h.x[4 * i + 3];
    set_skin(P, (h.contactmodel == 1 && h.db > 0.0) ? 0.0 : 0.1 * rmin);
    P.vl_cap = (uint32_t)NV;
}

// Place every simulation in the batch arrays: whole 256-particle chunks of
// particles, then walls, bonds, and a key range that holds the simulation's
// own keys (up to ncells + num[0] + num[0]*num[1]: a particle at exactly
// x == L or y == L survives the periodic wrap, P17) and its padding key.
void layout(sph_ctx* c) {
    uint32_t base = 0, walls = 0, bonds = 0, keys = 0, rh = 0, hrows = 0;
    bool shear = false, adaptive = false;
    for (size_t k = 0; k < c->sims.size(); ++k) {
        Sim& sim = c->sims[k];
        DevParams& P = sim.P;
        const Host& h = sim.h;
        P.sim = (uint32_t)k;
        P.base = base;
        P.npad = std::max<uint32_t>(BLOCK, (h.np + BLOCK - 1) / BLOCK * BLOCK);
        P.chunk_base = base / BLOCK;
        P.cell_base = keys;
        const uint32_t kmax = P.ncells + h.num[0] + h.num[0] * h.num[1];
        P.pad_key = keys + kmax + 1;
        P.wall_base = walls;
        P.bond_base = bonds;
        P.rhist_base = rh;
        P.hbase = hrows;
        hrows += h.np;
        base += P.npad;
        keys += kmax + 2;
        walls += h.nw;
        bonds += h.nb0;
        rh += 256u * n_tiles(P.npad);
        shear = shear || (h.nw > 0 && h.wmode[0] == 3);
        adaptive = adaptive || (h.adaptive == 1 && h.nw > 0);
    }
    DevArrays& A = c->A;
    A.nsim = (uint32_t)c->sims.size();
    A.ntot = base;
    A.nchunk = base / BLOCK;
    A.nbond = bonds;
    A.nkeys = keys;
    A.xcd_order = 0;
    c->sorted_bits = std::max(1, bit_length(keys));
    if (c->sorted_bits > 32) c->sorted_bits = 32;
    c->passes = (c->sorted_bits + c->radix_bits - 1) / c->radix_bits;
    c->shear_stress = shear;
    c->any_adaptive = adaptive;
    c->contactmodel = c->sims[0].h.contactmodel;
}

template <class T> int dev_alloc(sph_ctx* c, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) return hip_fail(c, e, "hipMalloc");
    c->allocs.push_back((void*)*p);
    c->alloc_bytes.push_back(count * sizeof(T));
    return SPH_OK;
}

#define ALLOC(ptr, n)                               \
    do {                                            \
        int rc_ = dev_alloc(c, &(ptr), (size_t)(n)); \
        if (rc_) return rc_;                         \
    } while (0)

int allocate(sph_ctx* c) {
    DevArrays& A = c->A;
    const size_t np = A.ntot, nb = A.nbond, nkeys = A.nkeys, ns = A.nsim;
    size_t nw = 0, rh = 0;
    for (const Sim& sim : c->sims) {
        nw += sim.h.nw;
        rh += 256u * n_tiles(sim.P.npad);
    }
    ALLOC(A.x, np); ALLOC(A.vel, np); ALLOC(A.angvel, np); ALLOC(A.angpos, np);
    ALLOC(A.xyzsum, np); ALLOC(A.acc, np); ALLOC(A.angacc, np);
    ALLOC(A.out, np * 4); ALLOC(A.es, np); ALLOC(A.ev, np);
    ALLOC(A.key, np); ALLOC(A.key_prev, np); ALLOC(A.skey, np); ALLOC(A.perm, np);
    ALLOC(A.key_tmp, np); ALLOC(A.val_tmp, np);
    // The batch-wide sort and the per-simulation one-workgroup sorts.
    ALLOC(A.rhist, std::max<size_t>(256 * (size_t)n_tiles((unsigned)np), rh));
    ALLOC(A.cell_start, nkeys); ALLOC(A.cell_end, nkeys);
    for (int k = 0; k < 2; ++k) {
        ALLOC(A.xs[k], np);
        ALLOC(A.vs[k], np);
        ALLOC(A.ws[k], np);
    }
    ALLOC(A.rank, np);
    ALLOC(A.vlist, np * NV);
    ALLOC(A.vcount, np);
    ALLOC(A.x_build, np);
    ALLOC(A.movers, ns * MV_CAP);
    ALLOC(A.skey_bk, np); ALLOC(A.perm_bk, np);
    ALLOC(A.cs_bk, nkeys); ALLOC(A.ce_bk, nkeys);
    ALLOC(A.crow, np * NC);
    ALLOC(A.perm_rows, np);
    ALLOC(A.delta_t, np * NC);
    ALLOC(A.w_nx, nw); ALLOC(A.w_mvfd, nw); ALLOC(A.w_acc, nw);
    ALLOC(A.tau_pp, np);
    ALLOC(A.partial, (size_t)MAXWALLS * A.nchunk);
    ALLOC(A.tau_partial, A.nchunk); ALLOC(A.tau_eff, ns);
    ALLOC(A.bonds, 2 * nb); ALLOC(A.bonds_delta, nb); ALLOC(A.bonds_omega, nb);
    ALLOC(A.bond_f, nb); ALLOC(A.bond_t, nb); ALLOC(A.bond_ok, nb);
    ALLOC(A.badj_off, np + 1);
    ALLOC(A.badj, 2 * nb);
    ALLOC(A.loop, ns);
    ALLOC(A.glob, 1);
    ALLOC(c->d_ptab, ns);
    A.ptab = c->d_ptab;
    uint32_t* chunk_sim = nullptr;
    uint32_t* bond_sim = nullptr;
    ALLOC(chunk_sim, A.nchunk);
    ALLOC(bond_sim, nb);
    A.chunk_sim = chunk_sim;
    A.bond_sim = bond_sim;
    return SPH_OK;
}

// The parameter table on the device (after a host-side parameter change).
int upload_params(sph_ctx* c) {
    std::vector<DevParams> tab(c->sims.size());
    for (size_t k = 0; k < c->sims.size(); ++k) tab[k] = c->sims[k].P;
    HIPCHK(c, hipMemcpyAsync(c->d_ptab, tab.data(), tab.size() * sizeof(DevParams),
                             hipMemcpyHostToDevice, c->stream), "H2D");
    // The copy reads `tab` asynchronously: wait before it goes out of scope.
    HIPCHK(c, hipStreamSynchronize(c->stream), "H2D");
    return SPH_OK;
}

// The output record of a particle (four double4 parts, OUT_* in
// dem_arrays.hpp) as 16 doubles: force xyz, 0 | torque xyz, 0 | es_dot,
// ev_dot, p, wall-0 force | wall 1-4 forces.
void pack_out(const Host& h, double* rec, size_t np) {
    for (size_t i = 0; i < np; ++i) {
        double* r = rec + 16 * i;
        r[0] = h.force[4 * i]; r[1] = h.force[4 * i + 1]; r[2] = h.force[4 * i + 2];
        r[4] = h.torque[4 * i]; r[5] = h.torque[4 * i + 1]; r[6] = h.torque[4 * i + 2];
        r[8] = h.es_dot[i]; r[9] = h.ev_dot[i]; r[10] = h.p[i];
    }
}

void unpack_out(const double* rec, Host& h, size_t np) {
    for (size_t i = 0; i < np; ++i) {
        const double* r = rec + 16 * i;
        h.force[4 * i] = r[0]; h.force[4 * i + 1] = r[1]; h.force[4 * i + 2] = r[2];
        h.torque[4 * i] = r[4]; h.torque[4 * i + 1] = r[5]; h.torque[4 * i + 2] = r[6];
        h.es_dot[i] = r[8]; h.ev_dot[i] = r[9]; h.p[i] = r[10];
    }
}

// Offset of a field in the 16 doubles of an output record, and its width
// (3 for vectors, 1 for scalars); -1 for fields not in the record.
int out_field(int f, int* width) {
    *width = 1;
    switch (f) {
        case SPH_F_FORCE: *width = 3; return 0;
        case SPH_F_TORQUE: *width = 3; return 4;
        case SPH_F_ES_DOT: return 8;
        case SPH_F_EV_DOT: return 9;
        case SPH_F_P: return 10;
        default: return -1;
    }
}

int upload_state(sph_ctx* c) {
    hipStream_t s = c->stream;
    DevArrays& A = c->A;
    const size_t N = A.ntot;
#define H2D(dst, src, bytes) \
    HIPCHK(c, hipMemcpyAsync((dst), (src), (bytes), hipMemcpyHostToDevice, s), "H2D")
    // Whole-batch initial values.  Padding particles keep zero state, their
    // padding key (never re-hashed) and empty rows.
    HIPCHK(c, hipMemsetAsync(A.x, 0, N * 32, s), "memset");
    HIPCHK(c, hipMemsetAsync(A.vel, 0, N * 32, s), "memset");
    HIPCHK(c, hipMemsetAsync(A.angvel, 0, N * 32, s), "memset");
    HIPCHK(c, hipMemsetAsync(A.angpos, 0, N * 32, s), "memset");
    HIPCHK(c, hipMemsetAsync(A.xyzsum, 0, N * 32, s), "memset");
    HIPCHK(c, hipMemsetAsync(A.acc, 0, N * 32, s), "memset");
    HIPCHK(c, hipMemsetAsync(A.angacc, 0, N * 32, s), "memset");
    HIPCHK(c, hipMemsetAsync(A.out, 0, N * 128, s), "memset");
    HIPCHK(c, hipMemsetAsync(A.es, 0, N * 8, s), "memset");
    HIPCHK(c, hipMemsetAsync(A.ev, 0, N * 8, s), "memset");
    HIPCHK(c, hipMemsetAsync(A.delta_t, 0, N * NC * sizeof(double4), s), "memset");
    HIPCHK(c, hipMemsetAsync(A.vcount, 0, N * 8, s), "memset");
    for (int k = 0; k < 2; ++k) {
        HIPCHK(c, hipMemsetAsync(A.xs[k], 0, N * 32, s), "memset");
        HIPCHK(c, hipMemsetAsync(A.vs[k], 0, N * 32, s), "memset");
        HIPCHK(c, hipMemsetAsync(A.ws[k], 0, N * 32, s), "memset");
    }
    HIPCHK(c, hipMemsetAsync(A.x_build, 0, N * 32, s), "memset");
    HIPCHK(c, hipMemsetAsync(A.tau_pp, 0, N * 8, s), "memset");
    HIPCHK(c, hipMemsetAsync(A.tau_eff, 0, A.nsim * 8, s), "memset");
    HIPCHK(c, hipMemsetAsync(A.glob, 0, sizeof(DevGlobal), s), "memset");
    // Contact rows start empty (the reference fills np, src/device.cu:458-462)
    // with zero displacement.
    std::vector<uint32_t> rows(N * NC, EMPTY);
    H2D(A.crow, rows.data(), N * NC * 4);
    std::vector<uint32_t> ident(N), keys(N), ff(N, 0xffffffffu), chunk_sim(A.nchunk);
    for (size_t i = 0; i < N; ++i) ident[i] = (uint32_t)i;
    H2D(A.perm, ident.data(), N * 4);
    H2D(A.rank, ident.data(), N * 4);
    H2D(A.perm_rows, ident.data(), N * 4);
    H2D(A.key_prev, ff.data(), N * 4);
    std::vector<uint32_t> cs(A.nkeys, 0xffffffffu);
    H2D(A.cell_start, cs.data(), (size_t)A.nkeys * 4);
    HIPCHK(c, hipMemsetAsync(A.cell_end, 0, (size_t)A.nkeys * 4, s), "memset");
    std::vector<double> wn, wm;
    std::vector<uint32_t> bonds, bond_sim;
    std::vector<double> bd, bo;
    std::vector<uint32_t> off(N + 1, 0u);
    std::vector<DevLoop> loops(A.nsim);
    std::vector<std::vector<double>> out_staging(c->sims.size());
    for (const Sim& sim : c->sims) {
        const Host& h = sim.h;
        const DevParams& P = sim.P;
        const size_t np = h.np, b0 = P.base;
        for (uint32_t i = 0; i < P.npad / BLOCK; ++i) chunk_sim[P.chunk_base + i] = P.sim;
        for (size_t i = 0; i < P.npad; ++i) keys[b0 + i] = P.pad_key;
        if (np) {
            H2D(A.x + b0, h.x.data(), np * 32);
            H2D(A.vel + b0, h.vel.data(), np * 32);
            H2D(A.angvel + b0, h.angvel.data(), np * 32);
            H2D(A.angpos + b0, h.angpos.data(), np * 32);
            H2D(A.xyzsum + b0, h.xyzsum.data(), np * 32);
            {
                std::vector<double>& rec = out_staging[P.sim];
                rec.assign(np * 16, 0.0);
                pack_out(h, rec.data(), np);
                H2D(A.out + 4 * b0, rec.data(), np * 128);
            }
            H2D(A.es + b0, h.es.data(), np * 8);
            H2D(A.ev + b0, h.ev.data(), np * 8);
            H2D(A.x_build + b0, h.x.data(), np * 32);
        }
        wn.insert(wn.end(), h.w_nx.begin(), h.w_nx.end());
        wm.insert(wm.end(), h.w_mvfd.begin(), h.w_mvfd.end());
        // Bonds in batch indices, with the adjacency (CSR, ascending bond id
        // per particle).
        const size_t nb = h.nb0;
        for (size_t b = 0; b < nb; ++b) {
            const uint32_t i = h.bonds[2 * b], j = h.bonds[2 * b + 1];
            bonds.push_back((uint32_t)(b0 + i));
            bonds.push_back((uint32_t)(b0 + j));
            bond_sim.push_back(P.sim);
            if (i < np) off[b0 + i + 1]++;
            if (j < np && j != i) off[b0 + j + 1]++;
        }
        bd.insert(bd.end(), h.bonds_delta.begin(), h.bonds_delta.end());
        bo.insert(bo.end(), h.bonds_omega.begin(), h.bonds_omega.end());
        DevLoop& L = loops[P.sim];
        memset(&L, 0, sizeof L);
        L.t = h.t;
        L.iter = 0;
        L.grid_Lz = h.L[2];
        L.velocity_state = 1;
        L.zero_empty = 1u;
        L.force_rebuild = 1u;
    }
    // Padding keys, and real keys before the first hash (the first step
    // always hashes).
    H2D(A.key, keys.data(), N * 4);
    H2D(A.skey, keys.data(), N * 4);
    H2D(const_cast<uint32_t*>(A.chunk_sim), chunk_sim.data(), chunk_sim.size() * 4);
    if (!wn.empty()) {
        H2D(A.w_nx, wn.data(), wn.size() * 8);
        H2D(A.w_mvfd, wm.data(), wm.size() * 8);
        HIPCHK(c, hipMemsetAsync(A.w_acc, 0, wn.size() / 4 * 8, s), "memset");
    }
    std::vector<uint32_t> adj;
    if (!bonds.empty()) {
        H2D(A.bonds, bonds.data(), bonds.size() * 4);
        H2D(const_cast<uint32_t*>(A.bond_sim), bond_sim.data(), bond_sim.size() * 4);
        H2D(A.bonds_delta, bd.data(), bd.size() * 8);
        H2D(A.bonds_omega, bo.data(), bo.size() * 8);
        for (size_t i = 0; i < N; ++i) off[i + 1] += off[i];
        adj.assign(off[N], 0u);
        std::vector<uint32_t> fillp(off.begin(), off.end() - 1);
        for (size_t b = 0; b < bond_sim.size(); ++b) {
            const Sim& sim = c->sims[bond_sim[b]];
            const uint32_t i = bonds[2 * b], j = bonds[2 * b + 1];
            const uint32_t b0 = sim.P.base, np = sim.h.np;
            if (i - b0 < np) adj[fillp[i]++] = (uint32_t)b;
            if (j - b0 < np && j != i) adj[fillp[j]++] = (uint32_t)b;
        }
        H2D(A.badj, adj.data(), adj.size() * 4);
    }
    H2D(A.badj_off, off.data(), (N + 1) * 4);
    H2D(A.loop, loops.data(), loops.size() * sizeof(DevLoop));
    c->hash_pending = true;
    c->prev_walls_resorted = false;
    for (Sim& sim : c->sims) sim.iter = 0;
#undef H2D
    HIPCHK(c, hipStreamSynchronize(s), "upload");
    return upload_params(c);
}

// readbin + checkValues for every image, then one set of device arrays for
// all of them.
// Device address of `scratch` on each device, looked up once per device
// with that device current (at the first context opened on it).
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[SCRATCH_MAX_DEVICES] = {};

hipError_t scratch_setup(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[device]) return hipSuccess;
    double* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[device] = ptr;
    return hipSuccess;
}

double* scratch_of(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_dev[device];
}

int open_images(const std::vector<std::pair<const void*, size_t>>& images, int device,
                unsigned flags, sph_ctx** out) {
    *out = nullptr;
    if (images.empty()) {
        g_open_error = "no simulations given";
        return SPH_E_ARG;
    }
    sph_ctx* c = new sph_ctx();
    c->flags = flags;
    c->sims.resize(images.size());
    uint64_t total = 0;
    for (size_t k = 0; k < images.size(); ++k) {
        Host& h = c->sims[k].h;
        std::string err;
        int rc = parse_bin(h, images[k].first, images[k].second, err);
        if (rc != SPH_OK) {
            g_open_error = images.size() > 1 ? "simulation " + std::to_string(k) + ": " + err : err;
            delete c;
            return rc;
        }
        for (uint32_t b = 0; b < h.nb0; ++b) {
            if (h.bonds[2 * b] >= h.np || h.bonds[2 * b + 1] >= h.np) {
                g_open_error = "Error: bond " + std::to_string(b) + " names a particle index >= np";
                delete c;
                return SPH_E_INPUT;
            }
        }
        total += std::max<uint32_t>(BLOCK, (h.np + BLOCK - 1) / BLOCK * BLOCK);
        if (!(flags & SPH_OPEN_NOCHECK)) {
            rc = check_values(c, h);
            if (rc != SPH_OK) {
                g_open_error = c->err;
                delete c;
                return rc;
            }
        }
        if (h.contactmodel != c->sims[0].h.contactmodel) {
            g_open_error = "all simulations of a batch must use the same contact model";
            delete c;
            return SPH_E_UNSUPPORTED;
        }
    }
    if (total >= MAX_NP) {
        g_open_error = "np exceeds the supported maximum (2^27)";
        delete c;
        return SPH_E_INPUT;
    }
    hipError_t e;
    if (device >= 0) {
        e = hipSetDevice(device);
        if (e != hipSuccess) {
            g_open_error = std::string("hipSetDevice: ") + hipGetErrorString(e);
            delete c;
            return SPH_E_DEVICE;
        }
    }
    (void)hipGetDevice(&c->device);
    e = scratch_setup(c->device);
    if (e != hipSuccess) {
        g_open_error = std::string("hipGetSymbolAddress: ") + hipGetErrorString(e);
        delete c;
        return SPH_E_DEVICE;
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        g_open_error = std::string("hipStreamCreate: ") + hipGetErrorString(e);
        delete c;
        return SPH_E_DEVICE;
    }
    memset(&c->A, 0, sizeof c->A);
    for (Sim& sim : c->sims) fill_params(sim);
    layout(c);
    if (hipHostMalloc((void**)&c->loop_sample, c->sims.size() * sizeof(DevLoop),
                      hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&c->sample_ev, hipEventDisableTiming) != hipSuccess) {
        // Without the sample the automatic placement keeps its first choice.
        (void)hipGetLastError();
        if (c->loop_sample) (void)hipHostFree(c->loop_sample);
        c->loop_sample = nullptr;
        c->sample_ev = nullptr;
    }
    int rc = allocate(c);
    if (rc == SPH_OK) rc = upload_state(c);
    if (rc != SPH_OK) {
        g_open_error = c->err;
        sph_close(c);
        return rc;
    }
    *out = c;
    return SPH_OK;
}

// -- step loop --------------------------------------------------------------

// Per-kernel timing with HIP events on the context stream.  Events are
// recorded around each launch without synchronising; durations are
// collected when the report is asked for.
struct ProfEvent {
    int stage;
    hipEvent_t e0, e1;
};

struct Prof {
    sph_ctx* c;
    int idx = -1;
    hipEvent_t e0 = nullptr;
    void begin(const char* name);
    void end();
};

}  // namespace

struct sph_prof_state {
    std::vector<ProfEvent> pending;
    std::vector<hipEvent_t> pool;
};

namespace {

sph_prof_state& prof_state(sph_ctx* c) {
    if (!c->prof) c->prof = new sph_prof_state();
    return *c->prof;
}

hipEvent_t take_event(sph_ctx* c) {
    sph_prof_state& ps = prof_state(c);
    if (!ps.pool.empty()) {
        hipEvent_t e = ps.pool.back();
        ps.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void Prof::begin(const char* name) {
    if (!c->profile) return;
    idx = -1;
    for (size_t k = 0; k < c->prof_names.size(); ++k)
        if (c->prof_names[k] == name) idx = (int)k;
    if (idx < 0) {
        c->prof_names.push_back(name);
        c->prof_ms.push_back(0.0);
        c->prof_n.push_back(0);
        idx = (int)c->prof_names.size() - 1;
    }
    e0 = take_event(c);
    (void)hipEventRecord(e0, c->stream);
}

void Prof::end() {
    if (!c->profile || idx < 0) return;
    hipEvent_t e1 = take_event(c);
    (void)hipEventRecord(e1, c->stream);
    prof_state(c).pending.push_back(ProfEvent{idx, e0, e1});
    idx = -1;
}

// Fold the recorded events into the per-stage totals.
void prof_collect(sph_ctx* c) {
    sph_prof_state& ps = prof_state(c);
    if (ps.pending.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (const ProfEvent& pe : ps.pending) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, pe.e0, pe.e1);
        c->prof_ms[pe.stage] += ms;
        c->prof_n[pe.stage] += 1;
        ps.pool.push_back(pe.e0);
        ps.pool.push_back(pe.e1);
    }
    ps.pending.clear();
}

// Re-sort placement for the steps about to be enqueued.  In automatic mode
// the choice follows the re-sort rate seen over the previous sph_step call,
// read from an asynchronous copy of the loop counters (no synchronisation):
// the end-of-step workgroup wins while fewer than one step in ten re-sorts.
void choose_sort_placement(sph_ctx* c) {
    if (c->sort_mode != 2) {
        c->resort_at_end = c->sort_mode == 1;
        return;
    }
    // A batch re-sorts each simulation in its own workgroup, in parallel,
    // which beats a batch-wide sort whenever any simulation changes.
    if (c->sims.size() > 1) {
        c->resort_at_end = true;
        return;
    }
    if (!c->sample_pending || !c->loop_sample) return;
    if (hipEventQuery(c->sample_ev) != hipSuccess) return;
    c->sample_pending = false;
    const DevLoop& L = c->loop_sample[0];
    const int64_t steps = L.iter - c->sample_iter0;
    const uint64_t sorts = L.sorts - c->sample_sorts0;
    if (steps >= 20) c->resort_at_end = (double)sorts < 0.1 * (double)steps;
    c->sample_iter0 = L.iter;
    c->sample_sorts0 = L.sorts;
}

void sample_loop(sph_ctx* c) {
    if (c->sort_mode != 2 || c->sims.size() > 1 || !c->loop_sample || c->sample_pending) return;
    if (hipMemcpyAsync(c->loop_sample, c->A.loop, sizeof(DevLoop), hipMemcpyDeviceToHost,
                       c->stream) != hipSuccess ||
        hipEventRecord(c->sample_ev, c->stream) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    c->sample_pending = true;
}

int enqueue_step(sph_ctx* c) {
    const DevArrays& A = c->A;
    hipStream_t s = c->stream;
    Prof pr{c};
    // With an adaptive grid L[2] is the top wall's position at the step
    // start, known only after the previous step's wall update.
    const bool hash_in_integrate = !c->any_adaptive;
    const bool hash_now = c->hash_pending || !hash_in_integrate;
    if (hash_now) {
        pr.begin("hash");
        HIPCHK(c, launch_hash(A, s), "hash");
        pr.end();
        c->hash_pending = false;
    }
    // The previous step's walls kernel already re-sorted for this step
    // unless this step hashes for itself.
    if (!(c->prev_walls_resorted && !hash_now)) {
        pr.begin("sort");
        HIPCHK(c, launch_sort(A, c->passes, c->radix_bits, s), "sort");
        pr.end();
        pr.begin("cells");
        HIPCHK(c, launch_cells(A, c->sorted_bits, s), "cells");
        pr.end();
    }
    // Contact and integration in one pass where each particle's force is
    // final after its own contact row: a fixed grid height, no bonds, no
    // wmode-3 wall, four lanes per particle.
    const bool fused = c->fused && hash_in_integrate && A.nbond == 0 && !c->shear_stress &&
                       c->contact_lanes == 4 && (c->contactmodel == 2 || c->contactmodel == 3);
    if (fused) {
        pr.begin("contact_integrate");
        HIPCHK(c, launch_contact_integrate(A, (int)c->contactmodel, s), "contact");
        pr.end();
        // Also after a re-sort: the order the fused pass left its rows in.
        pr.begin("wall_partials");
        HIPCHK(c, launch_wall_partials(A, s), "partials");
        pr.end();
    } else {
        pr.begin("contact");
        HIPCHK(c, launch_contact(A, (int)c->contactmodel, c->contact_lanes, c->contact_wg, s),
               "contact");
        pr.end();
    }
    if (!fused && A.nbond > 0) {
        pr.begin("bonds");
        HIPCHK(c, launch_bonds(A, s), "bonds");
        pr.end();
    }
    if (!fused && c->shear_stress) {
        pr.begin("shear_stress");
        HIPCHK(c, launch_shear_stress(A, s), "shear_stress");
        pr.end();
    }
    if (!fused && c->integrate_lanes == 4 && hash_in_integrate && !c->shear_stress) {
        pr.begin("integrate");
        HIPCHK(c, launch_integrate_q(A, s), "integrate");
        pr.end();
        bool walls = false;
        for (const Sim& sim : c->sims) walls = walls || sim.h.nw > 0;
        if (walls) {
            pr.begin("wall_partials");
            HIPCHK(c, launch_wall_partials(A, s), "partials");
            pr.end();
        }
    } else if (!fused) {
        pr.begin("integrate");
        HIPCHK(c, launch_integrate(A, hash_in_integrate ? 1 : 0, s), "integrate");
        pr.end();
    }
    const bool resort = hash_in_integrate && c->resort_at_end;
    pr.begin("walls");
    HIPCHK(c, launch_walls(A, resort ? 1 : 0, c->passes, c->radix_bits, c->sorted_bits, s),
           "walls");
    pr.end();
    c->prev_walls_resorted = resort;
    // Host mirror of the loop counters (src/device.cu:2536-2538).
    for (Sim& sim : c->sims) {
        sim.h.t += sim.h.dt;
        sim.iter += 1;
    }
    return SPH_OK;
}

void drop_graphs(sph_ctx* c) {
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
}

// Enqueue `n` steps as replays of a captured graph of c->graph_steps steps,
// capturing it first if needed.  Only in the steady state: no hash pending,
// no profiling, and the previous step's re-sort placement equal to the one
// the captured steps use (so the first captured step behaves like the rest).
// Returns the number of steps enqueued (0 when graphs do not apply).
int enqueue_graph_steps(sph_ctx* c, uint64_t n, uint64_t* done) {
    *done = 0;
    const int G = c->graph_steps;
    const bool resort = !c->any_adaptive && c->resort_at_end;
    if (G <= 0 || n < (uint64_t)G || c->profile || c->hash_pending ||
        c->prev_walls_resorted != resort)
        return SPH_OK;
    hipGraphExec_t exec = nullptr;
    for (auto& g : c->graphs)
        if (g.steps == G && g.resort == resort && g.lanes == c->contact_lanes) exec = g.exec;
    if (!exec) {
        // Capturing must not advance the host mirror of the loop.
        std::vector<std::pair<double, int64_t>> saved;
        for (const Sim& sim : c->sims) saved.push_back({sim.h.t, sim.iter});
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "capture");
        int rc = SPH_OK;
        for (int k = 0; k < G && rc == SPH_OK; ++k) rc = enqueue_step(c);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &graph);
        for (size_t k = 0; k < saved.size(); ++k) {
            c->sims[k].h.t = saved[k].first;
            c->sims[k].iter = saved[k].second;
        }
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(c, e, "capture");
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return hip_fail(c, e, "graph instantiate");
        c->graphs.push_back({G, resort, c->contact_lanes, exec});
    }
    while (n - *done >= (uint64_t)G) {
        HIPCHK(c, hipGraphLaunch(exec, c->stream), "graph launch");
        for (Sim& sim : c->sims)
            for (int k = 0; k < G; ++k) {
                sim.h.t += sim.h.dt;
                sim.iter += 1;
            }
        *done += (uint64_t)G;
    }
    return SPH_OK;
}

int read_loop(sph_ctx* c, int sim, DevLoop* L) {
    HIPCHK(c, hipMemcpyAsync(L, c->A.loop + sim, sizeof *L, hipMemcpyDeviceToHost, c->stream),
           "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "sync");
    return SPH_OK;
}

int check_device_flags(sph_ctx* c) {
    std::vector<DevLoop> L(c->sims.size());
    HIPCHK(c, hipMemcpyAsync(L.data(), c->A.loop, L.size() * sizeof(DevLoop),
                             hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "sync");
    for (size_t k = 0; k < L.size(); ++k)
        if (L[k].error_bits & DFLAG_KEY_RANGE)
            return fail(c, SPH_E_INPUT,
                        "%sOne or more particles moved outside the grid.\nThis could possibly "
                        "be caused by a numerical instability.\nIs the computational time step "
                        "too large?",
                        L.size() > 1 ? ("simulation " + std::to_string(k) + ": ").c_str() : "");
    return SPH_OK;
}

// D2H of all state the .bin carries, for simulation k.
int download_state(sph_ctx* c, int k) {
    Sim& sim = c->sims[k];
    Host& h = sim.h;
    const size_t np = h.np, nw = h.nw, nb = h.nb0;
    const size_t b0 = sim.P.base, w0 = sim.P.wall_base, bb = sim.P.bond_base;
    hipStream_t s = c->stream;
    const DevArrays& A = c->A;
    std::vector<double> out_rec;
#define D2H(dst, src, bytes) \
    HIPCHK(c, hipMemcpyAsync((dst), (src), (bytes), hipMemcpyDeviceToHost, s), "D2H")
    if (np) {
        D2H(h.x.data(), A.x + b0, np * 32);
        D2H(h.vel.data(), A.vel + b0, np * 32);
        D2H(h.angvel.data(), A.angvel + b0, np * 32);
        D2H(h.angpos.data(), A.angpos + b0, np * 32);
        D2H(h.xyzsum.data(), A.xyzsum + b0, np * 32);
        out_rec.resize(np * 16);
        D2H(out_rec.data(), A.out + 4 * b0, np * 128);
        D2H(h.es.data(), A.es + b0, np * 8);
        D2H(h.ev.data(), A.ev + b0, np * 8);
    }
    if (nw) {
        D2H(h.w_nx.data(), A.w_nx + w0, nw * 32);
        D2H(h.w_mvfd.data(), A.w_mvfd + w0, nw * 32);
    }
    if (nb) {
        D2H(h.bonds.data(), A.bonds + 2 * bb, nb * 8);
        D2H(h.bonds_delta.data(), A.bonds_delta + bb, nb * 32);
        D2H(h.bonds_omega.data(), A.bonds_omega + bb, nb * 32);
    }
#undef D2H
    DevLoop L;
    int rc = read_loop(c, k, &L);
    if (rc) return rc;
    for (size_t i = 0; i < 2 * nb; ++i) h.bonds[i] -= (uint32_t)b0;  // batch -> .bin index
    if (np) unpack_out(out_rec.data(), h, np);
    h.t = L.t;
    if (h.adaptive == 1 && h.nw > 0 && L.iter > 0) h.L[2] = L.grid_Lz;
    return SPH_OK;
}

int64_t field_bytes(const sph_ctx* c, int k, int f) {
    const Sim& sim = c->sims[k];
    const int64_t np = sim.h.np, nw = sim.h.nw, nb = sim.h.nb0;
    switch (f) {
        case SPH_F_X: case SPH_F_VEL: return np * 32;
        case SPH_F_ANGVEL: case SPH_F_ANGPOS: case SPH_F_FORCE: case SPH_F_TORQUE:
        case SPH_F_XYZSUM: case SPH_F_ACC: case SPH_F_ANGACC: return np * 24;
        case SPH_F_ES_DOT: case SPH_F_ES: case SPH_F_EV_DOT: case SPH_F_EV: case SPH_F_P:
            return np * 8;
        case SPH_F_CONTACTS: return np * NC * 4;
        case SPH_F_DELTA_T: return np * NC * 24;
        case SPH_F_DISTMOD: return np * NC * 32;
        case SPH_F_CELLID: case SPH_F_PERM: return np * 4;
        case SPH_F_CELLSTART: case SPH_F_CELLEND: return (int64_t)sim.P.ncells * 4;
        case SPH_F_WALLS_NX: case SPH_F_WALLS_MVFD: return nw * 32;
        case SPH_F_WALLS_ACC: return nw * 8;
        case SPH_F_WALLS_FORCE_PP: return nw * np * 8;
        case SPH_F_BONDS: return nb * 8;
        case SPH_F_BONDS_DELTA: case SPH_F_BONDS_OMEGA: return nb * 32;
        case SPH_F_COLOR: return np * 4;
        default: return -1;
    }
}

int d2h(sph_ctx* c, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SPH_OK;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "sync");
    return SPH_OK;
}
int h2d(sph_ctx* c, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SPH_OK;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream), "H2D");
    HIPCHK(c, hipStreamSynchronize(c->stream), "sync");
    return SPH_OK;
}

// double4 rows -> n x 3
void pack3(const std::vector<double>& v4, void* dst, size_t n) {
    double* d = static_cast<double*>(dst);
    for (size_t i = 0; i < n; ++i) {
        d[3 * i] = v4[4 * i]; d[3 * i + 1] = v4[4 * i + 1]; d[3 * i + 2] = v4[4 * i + 2];
    }
}

int get_vec3(sph_ctx* c, const double4* src, size_t n, void* dst) {
    std::vector<double> tmp(n * 4);
    int rc = d2h(c, tmp.data(), src, n * 32);
    if (rc) return rc;
    pack3(tmp, dst, n);
    return SPH_OK;
}

int set_vec3(sph_ctx* c, double4* dst, size_t n, const void* src) {
    std::vector<double> tmp(n * 4);
    int rc = d2h(c, tmp.data(), dst, n * 32);
    if (rc) return rc;
    const double* s = static_cast<const double*>(src);
    for (size_t i = 0; i < n; ++i) {
        tmp[4 * i] = s[3 * i]; tmp[4 * i + 1] = s[3 * i + 1]; tmp[4 * i + 2] = s[3 * i + 2];
    }
    return h2d(c, dst, tmp.data(), n * 32);
}

int mkdir_p(const std::string& path) {
    if (path.empty()) return 0;
    std::string cur;
    for (size_t i = 0; i < path.size(); ++i) {
        cur += path[i];
        if (path[i] == '/' || i + 1 == path.size()) mkdir(cur.c_str(), 0755);
    }
    return 0;
}

int write_status(sph_ctx* c, const std::string& fn) {
    FILE* fp = fopen(fn.c_str(), "w");
    if (!fp) return fail(c, SPH_E_IO, "could not write %s", fn.c_str());
    const Host& h = c->h0();
    fprintf(fp, "%2.4e %2.4e %d\n", h.t, 100.0 * h.t / h.t_total, h.step_count);
    fclose(fp);
    return SPH_OK;
}

}  // namespace

// ===========================================================================
namespace {
// "<dir>/<name>.bin" -> "<dir>/<name>.state"
std::string state_path(const std::string& bin) {
    const size_t n = bin.size();
    if (n >= 4 && bin.compare(n - 4, 4, ".bin") == 0) return bin.substr(0, n - 4) + ".state";
    return bin + ".state";
}
}  // namespace

extern "C" {

int sph_open_mem(const void* bytes, size_t nbytes, int device, unsigned flags, sph_ctx** out) {
    if (!out || (!bytes && nbytes)) return SPH_E_ARG;
    return open_images({{bytes, nbytes}}, device, flags, out);
}

namespace {
int read_file(const char* path, std::vector<unsigned char>& buf) {
    FILE* fp = fopen(path, "rb");
    if (!fp) {
        g_open_error = std::string("Could not read input binary file '") + path + "'";
        return SPH_E_IO;
    }
    unsigned char tmp[1 << 16];
    size_t k;
    while ((k = fread(tmp, 1, sizeof tmp, fp)) > 0) buf.insert(buf.end(), tmp, tmp + k);
    fclose(fp);
    return SPH_OK;
}
}  // namespace

int sph_open(const char* bin_path, int device, unsigned flags, sph_ctx** out) {
    if (!out || !bin_path) return SPH_E_ARG;
    *out = nullptr;
    std::vector<unsigned char> buf;
    int rc = read_file(bin_path, buf);
    if (rc) return rc;
    rc = open_images({{buf.data(), buf.size()}}, device, flags, out);
    if (rc || !(flags & SPH_OPEN_SIDECAR)) return rc;
    const std::string sp = state_path(bin_path);
    FILE* fp = fopen(sp.c_str(), "rb");
    if (!fp) return SPH_OK;  // no sidecar: the reference's fresh start
    fclose(fp);
    rc = sph_read_state(*out, sp.c_str());
    if (rc) {
        g_open_error = sph_last_error(*out);
        sph_close(*out);
        *out = nullptr;
    }
    return rc;
}

int sph_open_batch_mem(const void* const* images, const size_t* sizes, int nsims, int device,
                       unsigned flags, sph_ctx** out) {
    if (!out || !images || !sizes || nsims <= 0) return SPH_E_ARG;
    std::vector<std::pair<const void*, size_t>> v;
    for (int k = 0; k < nsims; ++k) {
        if (!images[k] && sizes[k]) return SPH_E_ARG;
        v.push_back({images[k], sizes[k]});
    }
    return open_images(v, device, flags, out);
}

int sph_open_batch(const char* const* bin_paths, int nsims, int device, unsigned flags,
                   sph_ctx** out) {
    if (!out || !bin_paths || nsims <= 0) return SPH_E_ARG;
    *out = nullptr;
    std::vector<std::vector<unsigned char>> bufs(nsims);
    std::vector<std::pair<const void*, size_t>> v;
    for (int k = 0; k < nsims; ++k) {
        if (!bin_paths[k]) return SPH_E_ARG;
        int rc = read_file(bin_paths[k], bufs[k]);
        if (rc) return rc;
        v.push_back({bufs[k].data(), bufs[k].size()});
    }
    return open_images(v, device, flags, out);
}

int sph_num_sims(const sph_ctx* c) { return c ? (int)c->sims.size() : SPH_E_ARG; }

int sph_step(sph_ctx* c, uint64_t nsteps) {
    if (!c) return SPH_E_ARG;
    choose_sort_placement(c);
    uint64_t k = 0;
    while (k < nsteps) {
        uint64_t done = 0;
        int rc = enqueue_graph_steps(c, nsteps - k, &done);
        if (rc) return rc;
        if (done == 0) {
            rc = enqueue_step(c);
            if (rc) return rc;
            done = 1;
        }
        k += done;
    }
    sample_loop(c);
    return SPH_OK;
}

int sph_sync(sph_ctx* c) {
    if (!c) return SPH_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream), "sync");
    return check_device_flags(c);
}

#define CHECK_SIM(c, sim) \
    do { if (!(c) || (sim) < 0 || (size_t)(sim) >= (c)->sims.size()) return SPH_E_ARG; } while (0)

int sph_write_mem_sim(sph_ctx* c, int sim, void* dst, size_t cap, size_t* nbytes) {
    CHECK_SIM(c, sim);
    int rc = download_state(c, sim);
    if (rc) return rc;
    Writer w;
    write_bin(c->sims[sim].h, w);
    if (nbytes) *nbytes = w.b.size();
    if (dst && cap >= w.b.size()) memcpy(dst, w.b.data(), w.b.size());
    return SPH_OK;
}

int sph_write_mem(sph_ctx* c, void* dst, size_t cap, size_t* nbytes) {
    return sph_write_mem_sim(c, 0, dst, cap, nbytes);
}

int sph_write_bin_sim(sph_ctx* c, int sim, const char* path) {
    CHECK_SIM(c, sim);
    if (!path) return SPH_E_ARG;
    int rc = download_state(c, sim);
    if (rc) return rc;
    Writer w;
    write_bin(c->sims[sim].h, w);
    FILE* fp = fopen(path, "wb");
    if (!fp) return fail(c, SPH_E_IO, "could create output binary file '%s'", path);
    const size_t k = fwrite(w.b.data(), 1, w.b.size(), fp);
    fclose(fp);
    if (k != w.b.size()) return fail(c, SPH_E_IO, "short write to '%s'", path);
    return SPH_OK;
}

int sph_write_bin(sph_ctx* c, const char* path) { return sph_write_bin_sim(c, 0, path); }

// Restart sidecar: the device state a .bin does not carry (src/file_io.cpp
// writes none of it, so the reference restarts with empty contact rows,
// zero displacements and acc0 = 0).  Little-endian, no padding:
//   "SPHSTATE" | u32 version (1) | u32 np | u32 nw | u32 NC | i64 iter |
//   u32 partner[np*NC] (np = empty) | u8 image[np*NC] (x code | y code << 2,
//   codes 0: none, 1: +L, 2: -L) | f64 delta_t[np*NC*3] | f64 acc[np*3] |
//   f64 angacc[np*3] | f64 walls_acc[nw]
namespace {
// A simulation's contact rows with the partners' original (batch) indices.
// The device words index the order of perm_rows (the sorted order of the
// step that wrote them).
int rows_get(sph_ctx* c, int sim, std::vector<uint32_t>& rows) {
    const Sim& S = c->sims[sim];
    const size_t np = S.h.np, b0 = S.P.base, npad = S.P.npad;
    rows.resize(np * NC);
    std::vector<uint32_t> pr(npad);
    int rc = d2h(c, rows.data(), c->A.crow + b0 * NC, np * NC * 4);
    if (!rc) rc = d2h(c, pr.data(), c->A.perm_rows + b0, npad * 4);
    if (rc) return rc;
    for (uint32_t& w : rows) {
        const uint32_t idx = w & CW_IDX_MASK;
        if (idx != EMPTY && idx - b0 < npad) w = (w & ~CW_IDX_MASK) | pr[idx - b0];
    }
    return SPH_OK;
}

int set_loop_flag(sph_ctx* c, int sim, uint32_t DevLoop::*flag);

// Writes rows given with original (batch) indices, in the order of
// perm_rows, and has the next step rebuild the candidate lists (their hit
// masks describe the rows they saw).
int rows_put(sph_ctx* c, int sim, std::vector<uint32_t>& rows) {
    const Sim& S = c->sims[sim];
    const size_t np = S.h.np, b0 = S.P.base, npad = S.P.npad;
    std::vector<uint32_t> pr(npad), inv(npad);
    int rc = d2h(c, pr.data(), c->A.perm_rows + b0, npad * 4);
    if (rc) return rc;
    for (size_t k = 0; k < npad; ++k) inv[pr[k] - b0] = (uint32_t)(b0 + k);
    for (uint32_t& w : rows) {
        const uint32_t idx = w & CW_IDX_MASK;
        if (idx != EMPTY && idx - b0 < npad) w = (w & ~CW_IDX_MASK) | inv[idx - b0];
    }
    rc = h2d(c, c->A.crow + b0 * NC, rows.data(), np * NC * 4);
    if (!rc) rc = set_loop_flag(c, sim, &DevLoop::force_rebuild);
    return rc;
}
}  // namespace

namespace {
constexpr char STATE_MAGIC[8] = {'S', 'P', 'H', 'S', 'T', 'A', 'T', 'E'};
constexpr uint32_t STATE_VERSION = 1;
}  // namespace

int sph_write_state_sim(sph_ctx* c, int sim, const char* path) {
    CHECK_SIM(c, sim);
    if (!path) return SPH_E_ARG;
    const Sim& S = c->sims[sim];
    const size_t np = S.h.np, nw = S.h.nw, b0 = S.P.base;
    DevLoop L;
    int rc = read_loop(c, sim, &L);
    std::vector<uint32_t> rows;
    std::vector<double> dt(np * NC * 3), acc(np * 3), angacc(np * 3), wacc(nw);
    if (!rc) rc = rows_get(c, sim, rows);
    if (!rc) rc = sph_get_field_sim(c, sim, SPH_F_DELTA_T, dt.data(), dt.size() * 8);
    if (!rc) rc = sph_get_field_sim(c, sim, SPH_F_ACC, acc.data(), acc.size() * 8);
    if (!rc) rc = sph_get_field_sim(c, sim, SPH_F_ANGACC, angacc.data(), angacc.size() * 8);
    if (!rc && nw) rc = sph_get_field_sim(c, sim, SPH_F_WALLS_ACC, wacc.data(), nw * 8);
    if (rc) return rc;
    Writer w;
    w.b.insert(w.b.end(), STATE_MAGIC, STATE_MAGIC + 8);
    w.put(STATE_VERSION);
    w.put((uint32_t)np);
    w.put((uint32_t)nw);
    w.put((uint32_t)NC);
    w.put((int64_t)L.iter);
    for (uint32_t v : rows) {
        const uint32_t idx = v & CW_IDX_MASK;
        w.put(idx == EMPTY ? (uint32_t)np : idx - (uint32_t)b0);
    }
    for (uint32_t v : rows) {
        const uint32_t cx = (v >> CW_SHIFT_X) & 3u, cy = (v >> CW_SHIFT_Y) & 3u;
        w.b.push_back((char)(cx | (cy << 2)));
    }
    for (double v : dt) w.put(v);
    for (double v : acc) w.put(v);
    for (double v : angacc) w.put(v);
    for (double v : wacc) w.put(v);
    FILE* fp = fopen(path, "wb");
    if (!fp) return fail(c, SPH_E_IO, "could not create state file '%s'", path);
    const size_t k = fwrite(w.b.data(), 1, w.b.size(), fp);
    fclose(fp);
    if (k != w.b.size()) return fail(c, SPH_E_IO, "short write to '%s'", path);
    return SPH_OK;
}

int sph_write_state(sph_ctx* c, const char* path) { return sph_write_state_sim(c, 0, path); }

int sph_read_state_sim(sph_ctx* c, int sim, const char* path) {
    CHECK_SIM(c, sim);
    if (!path) return SPH_E_ARG;
    Sim& S = c->sims[sim];
    const size_t np = S.h.np, nw = S.h.nw, b0 = S.P.base;
    FILE* fp = fopen(path, "rb");
    if (!fp) return fail(c, SPH_E_IO, "could not open state file '%s'", path);
    std::vector<char> buf;
    char tmp[1 << 16];
    size_t k;
    while ((k = fread(tmp, 1, sizeof tmp, fp)) > 0) buf.insert(buf.end(), tmp, tmp + k);
    fclose(fp);
    const size_t need = 8 + 4 * 4 + 8 + np * NC * (4 + 1 + 24) + np * 48 + nw * 8;
    if (buf.size() < 28 || memcmp(buf.data(), STATE_MAGIC, 8) != 0)
        return fail(c, SPH_E_INPUT, "'%s' is not a state file", path);
    uint32_t hdr[4];
    memcpy(hdr, buf.data() + 8, 16);
    if (hdr[0] != STATE_VERSION)
        return fail(c, SPH_E_INPUT, "'%s': state file version %u (expected %u)", path, hdr[0],
                    STATE_VERSION);
    if (hdr[1] != np || hdr[2] != nw || hdr[3] != (uint32_t)NC)
        return fail(c, SPH_E_INPUT,
                    "'%s' holds np %u, nw %u, NC %u; the simulation has np %zu, nw %zu, NC %d",
                    path, hdr[1], hdr[2], hdr[3], np, nw, NC);
    if (buf.size() != need)
        return fail(c, SPH_E_INPUT, "'%s': %zu bytes, expected %zu", path, buf.size(), need);
    const char* q = buf.data() + 24;
    int64_t iter;
    memcpy(&iter, q, 8);
    q += 8;
    if (iter < 0) return fail(c, SPH_E_INPUT, "'%s': negative iteration count", path);
    std::vector<uint32_t> rows(np * NC);
    for (size_t i = 0; i < np * NC; ++i) {
        uint32_t v;
        memcpy(&v, q + 4 * i, 4);
        if (v > np) return fail(c, SPH_E_INPUT, "'%s': contact index out of range", path);
        rows[i] = v == np ? EMPTY : v + (uint32_t)b0;
    }
    q += np * NC * 4;
    for (size_t i = 0; i < np * NC; ++i) {
        const uint32_t img = (uint8_t)q[i], cx = img & 3u, cy = (img >> 2) & 3u;
        if (cx == 3u || cy == 3u || img > 15u)
            return fail(c, SPH_E_INPUT, "'%s': bad image code", path);
        if (rows[i] != EMPTY) rows[i] |= (cx << CW_SHIFT_X) | (cy << CW_SHIFT_Y);
    }
    q += np * NC;
    std::vector<double> dt(np * NC * 3), acc(np * 3), angacc(np * 3), wacc(nw);
    memcpy(dt.data(), q, dt.size() * 8);
    q += dt.size() * 8;
    memcpy(acc.data(), q, acc.size() * 8);
    q += acc.size() * 8;
    memcpy(angacc.data(), q, angacc.size() * 8);
    q += angacc.size() * 8;
    memcpy(wacc.data(), q, nw * 8);
    int rc = rows_put(c, sim, rows);
    if (!rc) rc = sph_set_field_sim(c, sim, SPH_F_DELTA_T, dt.data(), dt.size() * 8);
    if (!rc) rc = sph_set_field_sim(c, sim, SPH_F_ACC, acc.data(), acc.size() * 8);
    if (!rc) rc = sph_set_field_sim(c, sim, SPH_F_ANGACC, angacc.data(), angacc.size() * 8);
    if (!rc && nw) rc = sph_set_field_sim(c, sim, SPH_F_WALLS_ACC, wacc.data(), nw * 8);
    if (!rc) rc = sph_set_scalar_sim(c, sim, SPH_S_ITER, (double)iter);
    if (!rc) rc = set_loop_flag(c, sim, &DevLoop::force_rebuild);
    return rc;
}

int sph_read_state(sph_ctx* c, const char* path) { return sph_read_state_sim(c, 0, path); }

// An output .bin, and with SPH_OPEN_SIDECAR its restart sidecar.
static int write_output(sph_ctx* c, const char* file) {
    int rc = sph_write_bin(c, file);
    if (rc || !(c->flags & SPH_OPEN_SIDECAR)) return rc;
    return sph_write_state(c, state_path(file).c_str());
}

int sph_run(sph_ctx* c, const char* out_dir, const char* sid) {
    if (!c || !out_dir || !sid) return SPH_E_ARG;
    if (c->sims.size() != 1)
        return fail(c, SPH_E_UNSUPPORTED, "sph_run drives one simulation; step a batch with "
                    "sph_step and write each simulation with sph_write_bin_sim");
    Host& h = c->h0();
    const std::string dir(out_dir);
    mkdir_p(dir);
    const bool verbose = !(c->flags & SPH_OPEN_QUIET);
    char file[4096];
    int rc;
    HIPCHK(c, hipStreamSynchronize(c->stream), "sync");
    if (h.step_count == 0) {
        snprintf(file, sizeof file, "%s/%s.output00000.bin", dir.c_str(), sid);
        if ((rc = write_output(c, file))) return rc;
    }
    double filetimeclock = 0.0;
    const std::string status = dir + "/" + sid + ".status.dat";
    if ((rc = write_status(c, status))) return rc;
    const double t_start = h.t;
    const clock_t tic = clock();
    long iter_run = 0;
    while (h.t <= h.t_total) {
        // Steps until the next output or the end, computed with the same
        // double additions the reference's host loop performs.
        uint64_t k = 0;
        double t = h.t, ftc = filetimeclock;
        bool out = false;
        while (t <= h.t_total) {
            t += h.dt;
            ftc += h.dt;
            ++k;
            if (ftc >= h.file_dt) { out = true; break; }
            if (k >= 1000) break;
        }
        const double t_before = h.t;
        if ((rc = sph_step(c, k))) return rc;
        // sph_step advanced h.t by the same additions.
        (void)t_before;
        filetimeclock = ftc;
        iter_run += (long)k;
        if (verbose && (iter_run % 10 == 0 || out)) {
            const double spent = (double)(clock() - tic) / CLOCKS_PER_SEC;
            const double ratio = spent / (h.t - t_start);
            printf("\r  Current time: %g/%g s. (%g s_real/s_sim)       ", h.t, h.t_total, ratio);
            fflush(stdout);
        }
        if (out) {
            if ((rc = sph_sync(c))) return rc;
            h.step_count += 1;
            snprintf(file, sizeof file, "%s/%s.output%05d.bin", dir.c_str(), sid, h.step_count);
            if ((rc = write_output(c, file))) return rc;
            if ((rc = write_status(c, status))) return rc;
            filetimeclock = 0.0;
        }
    }
    if ((rc = sph_sync(c))) return rc;
    if (verbose) printf("\nSimulation ended. Last output file number: %u\n", h.step_count);
    return SPH_OK;
}

int64_t sph_field_bytes_sim(sph_ctx* c, int sim, int field) {
    CHECK_SIM(c, sim);
    return field_bytes(c, sim, field);
}

int64_t sph_field_bytes(sph_ctx* c, int field) { return sph_field_bytes_sim(c, 0, field); }

int sph_get_field_sim(sph_ctx* c, int sim, int f, void* dst, size_t bytes) {
    CHECK_SIM(c, sim);
    const int64_t need = field_bytes(c, sim, f);
    if (need < 0 || (size_t)need != bytes || (!dst && bytes))
        return fail(c, SPH_E_ARG, "field %d: expected %lld bytes", f, (long long)need);
    const Sim& S = c->sims[sim];
    const size_t np = S.h.np, b0 = S.P.base, w0 = S.P.wall_base, bb = S.P.bond_base;
    const DevArrays& A = c->A;
    switch (f) {
        case SPH_F_X: return d2h(c, dst, A.x + b0, bytes);
        case SPH_F_VEL: return d2h(c, dst, A.vel + b0, bytes);
        case SPH_F_ANGVEL: return get_vec3(c, A.angvel + b0, np, dst);
        case SPH_F_ANGPOS: return get_vec3(c, A.angpos + b0, np, dst);

        case SPH_F_XYZSUM: return get_vec3(c, A.xyzsum + b0, np, dst);
        case SPH_F_ACC: return get_vec3(c, A.acc + b0, np, dst);
        case SPH_F_ANGACC: return get_vec3(c, A.angacc + b0, np, dst);
        case SPH_F_ES: return d2h(c, dst, A.es + b0, bytes);
        case SPH_F_EV: return d2h(c, dst, A.ev + b0, bytes);
        case SPH_F_FORCE:
        case SPH_F_TORQUE:
        case SPH_F_ES_DOT:
        case SPH_F_EV_DOT:
        case SPH_F_P: {
            // From the output records.
            int width;
            const int off = out_field(f, &width);
            std::vector<double> rec(np * 16);
            int rc = d2h(c, rec.data(), A.out + 4 * b0, np * 128);
            if (rc) return rc;
            double* d = static_cast<double*>(dst);
            for (size_t i = 0; i < np; ++i)
                for (int k = 0; k < width; ++k) d[width * i + k] = rec[16 * i + off + k];
            return SPH_OK;
        }
        case SPH_F_CONTACTS: {
            // The reference's row format: partner's .bin index, np when empty.
            std::vector<uint32_t> rows;
            int rc = rows_get(c, sim, rows);
            if (rc) return rc;
            uint32_t* w = static_cast<uint32_t*>(dst);
            for (size_t i = 0; i < np * NC; ++i) {
                const uint32_t idx = rows[i] & CW_IDX_MASK;
                w[i] = idx == EMPTY ? (uint32_t)np : idx - (uint32_t)b0;
            }
            return SPH_OK;
        }
        case SPH_F_DELTA_T: return get_vec3(c, A.delta_t + b0 * NC, np * NC, dst);
        case SPH_F_DISTMOD: {
            // Reconstructed from the packed image code and the partner's
            // radius; meaningful where the slot is occupied.
            std::vector<uint32_t> rows;
            std::vector<double> x(np * 4);
            int rc = rows_get(c, sim, rows);
            if (!rc) rc = d2h(c, x.data(), A.x + b0, np * 32);
            if (rc) return rc;
            double* d = static_cast<double*>(dst);
            for (size_t k = 0; k < np * NC; ++k) {
                const uint32_t wv = rows[k], idx = (wv & CW_IDX_MASK) - (uint32_t)b0;
                const uint32_t cx = (wv >> CW_SHIFT_X) & 3u, cy = (wv >> CW_SHIFT_Y) & 3u;
                d[4 * k] = cx == 1 ? S.h.L[0] : cx == 2 ? -S.h.L[0] : 0.0;
                d[4 * k + 1] = cy == 1 ? S.h.L[1] : cy == 2 ? -S.h.L[1] : 0.0;
                d[4 * k + 2] = 0.0;
                d[4 * k + 3] = idx < np ? x[4 * idx + 3] : 0.0;
            }
            return SPH_OK;
        }
        case SPH_F_CELLID:
        case SPH_F_PERM:
        case SPH_F_CELLSTART:
        case SPH_F_CELLEND: {
            // After a step that ended by re-sorting for the next one, the
            // order the step itself used is in the backup arrays.
            DevLoop L;
            int rc = read_loop(c, sim, &L);
            if (rc) return rc;
            const bool bk = L.resorted_last != 0u;
            const uint32_t* src = nullptr;
            size_t off = 0;
            if (f == SPH_F_CELLID) { src = bk ? A.skey_bk : A.skey; off = b0; }
            if (f == SPH_F_PERM) { src = bk ? A.perm_bk : A.perm; off = b0; }
            if (f == SPH_F_CELLSTART) { src = bk ? A.cs_bk : A.cell_start; off = S.P.cell_base; }
            if (f == SPH_F_CELLEND) { src = bk ? A.ce_bk : A.cell_end; off = S.P.cell_base; }
            rc = d2h(c, dst, src + off, bytes);
            uint32_t* w = static_cast<uint32_t*>(dst);
            if (f == SPH_F_CELLID)
                for (size_t i = 0; !rc && i < np; ++i) w[i] -= S.P.cell_base;
            if (f == SPH_F_PERM)
                for (size_t i = 0; !rc && i < np; ++i) w[i] -= (uint32_t)b0;
            if (f == SPH_F_CELLSTART || f == SPH_F_CELLEND)
                for (size_t i = 0; !rc && i < S.P.ncells; ++i)
                    if (w[i] != 0xffffffffu) w[i] -= (uint32_t)b0;
            return rc;
        }
        case SPH_F_WALLS_NX: return d2h(c, dst, A.w_nx + w0, bytes);
        case SPH_F_WALLS_MVFD: return d2h(c, dst, A.w_mvfd + w0, bytes);
        case SPH_F_WALLS_ACC: return d2h(c, dst, A.w_acc + w0, bytes);
        case SPH_F_WALLS_FORCE_PP: {
            // Wall-major, from the output records (doubles 11 and 12..15).
            std::vector<double> rec(np * 16);
            int rc = np ? d2h(c, rec.data(), A.out + 4 * b0, np * 128) : SPH_OK;
            if (rc) return rc;
            double* d = static_cast<double*>(dst);
            for (uint32_t w = 0; w < S.h.nw; ++w)
                for (size_t i = 0; i < np; ++i)
                    d[(size_t)w * np + i] = w < OUT_NWALL ? rec[16 * i + 11 + w] : 0.0;
            return SPH_OK;
        }
        case SPH_F_BONDS: {
            int rc = d2h(c, dst, A.bonds + 2 * bb, bytes);
            uint32_t* w = static_cast<uint32_t*>(dst);
            for (size_t i = 0; !rc && i < 2 * (size_t)S.h.nb0; ++i) w[i] -= (uint32_t)b0;
            return rc;
        }
        case SPH_F_BONDS_DELTA: return d2h(c, dst, A.bonds_delta + bb, bytes);
        case SPH_F_BONDS_OMEGA: return d2h(c, dst, A.bonds_omega + bb, bytes);
        case SPH_F_COLOR:
            if (bytes) memcpy(dst, S.h.color.data(), bytes);
            return SPH_OK;
        default: return fail(c, SPH_E_ARG, "unknown field %d", f);
    }
}

int sph_get_field(sph_ctx* c, int f, void* dst, size_t bytes) {
    return sph_get_field_sim(c, 0, f, dst, bytes);
}

namespace {
int set_loop_flag(sph_ctx* c, int sim, uint32_t DevLoop::*flag) {
    const uint32_t one = 1u;
    return h2d(c, &(c->A.loop[sim].*flag), &one, 4);
}
}  // namespace

int sph_set_field_sim(sph_ctx* c, int sim, int f, const void* src, size_t bytes) {
    CHECK_SIM(c, sim);
    const int64_t need = field_bytes(c, sim, f);
    if (need < 0 || (size_t)need != bytes || (!src && bytes))
        return fail(c, SPH_E_ARG, "field %d: expected %lld bytes", f, (long long)need);
    Sim& S = c->sims[sim];
    const size_t np = S.h.np, b0 = S.P.base, w0 = S.P.wall_base, bb = S.P.bond_base;
    DevArrays& A = c->A;
    switch (f) {
        case SPH_F_X:
        case SPH_F_VEL:
        case SPH_F_ANGVEL: {
            // New positions or velocities: the next step hashes and re-sorts
            // (refreshing the sorted copies the contact search reads), and
            // new positions rebuild the candidate lists.
            int rc = f == SPH_F_X     ? h2d(c, A.x + b0, src, bytes)
                     : f == SPH_F_VEL ? h2d(c, A.vel + b0, src, bytes)
                                      : set_vec3(c, A.angvel + b0, np, src);
            if (rc) return rc;
            HIPCHK(c, hipMemsetAsync(A.key_prev + b0, 0xff, np * 4, c->stream), "memset");
            c->hash_pending = true;
            return f == SPH_F_X ? set_loop_flag(c, sim, &DevLoop::force_rebuild) : SPH_OK;
        }
        case SPH_F_ANGPOS: return set_vec3(c, A.angpos + b0, np, src);

        case SPH_F_XYZSUM: return set_vec3(c, A.xyzsum + b0, np, src);
        case SPH_F_ACC: return set_vec3(c, A.acc + b0, np, src);
        case SPH_F_ANGACC: return set_vec3(c, A.angacc + b0, np, src);
        case SPH_F_ES: return h2d(c, A.es + b0, src, bytes);
        case SPH_F_EV: return h2d(c, A.ev + b0, src, bytes);
        case SPH_F_FORCE:
        case SPH_F_TORQUE:
        case SPH_F_ES_DOT:
        case SPH_F_EV_DOT:
        case SPH_F_P: {
            // Into the output records, keeping their other fields.
            int width;
            const int off = out_field(f, &width);
            std::vector<double> rec(np * 16);
            int rc = d2h(c, rec.data(), A.out + 4 * b0, np * 128);
            if (rc) return rc;
            const double* v = static_cast<const double*>(src);
            for (size_t i = 0; i < np; ++i)
                for (int k = 0; k < width; ++k) rec[16 * i + off + k] = v[width * i + k];
            return h2d(c, A.out + 4 * b0, rec.data(), np * 128);
        }
        case SPH_F_CONTACTS: {
            // Keep each slot's image code; replace the partner index.
            std::vector<uint32_t> rows;
            int rc = rows_get(c, sim, rows);
            if (rc) return rc;
            const uint32_t* s = static_cast<const uint32_t*>(src);
            for (size_t k = 0; k < np * NC; ++k) {
                if (s[k] > np) return fail(c, SPH_E_ARG, "contact index out of range");
                const uint32_t idx = s[k] == np ? EMPTY : s[k] + (uint32_t)b0;
                rows[k] = (rows[k] & ~CW_IDX_MASK) | idx;
            }
            rc = rows_put(c, sim, rows);
            if (rc) return rc;
            return set_loop_flag(c, sim, &DevLoop::zero_empty);
        }
        case SPH_F_DELTA_T: {
            int rc = set_vec3(c, A.delta_t + b0 * NC, np * NC, src);
            if (rc) return rc;
            return set_loop_flag(c, sim, &DevLoop::zero_empty);
        }
        case SPH_F_DISTMOD: {
            std::vector<uint32_t> rows;
            int rc = rows_get(c, sim, rows);
            if (rc) return rc;
            const double* d = static_cast<const double*>(src);
            for (size_t k = 0; k < np * NC; ++k) {
                uint32_t code = 0;
                const double sx = d[4 * k], sy = d[4 * k + 1];
                if (sx == S.h.L[0]) code |= 1u << CW_SHIFT_X;
                else if (sx == -S.h.L[0]) code |= 2u << CW_SHIFT_X;
                else if (sx != 0.0) return fail(c, SPH_E_ARG, "distmod x must be 0 or +-L[0]");
                if (sy == S.h.L[1]) code |= 1u << CW_SHIFT_Y;
                else if (sy == -S.h.L[1]) code |= 2u << CW_SHIFT_Y;
                else if (sy != 0.0) return fail(c, SPH_E_ARG, "distmod y must be 0 or +-L[1]");
                if (d[4 * k + 2] != 0.0) return fail(c, SPH_E_ARG, "distmod z must be 0");
                rows[k] = (rows[k] & CW_IDX_MASK) | code;
            }
            return rows_put(c, sim, rows);
        }
        case SPH_F_WALLS_NX: {
            int rc = h2d(c, A.w_nx + w0, src, bytes);
            if (!rc && bytes) memcpy(S.h.w_nx.data(), src, bytes);
            return rc;
        }
        case SPH_F_WALLS_MVFD: {
            int rc = h2d(c, A.w_mvfd + w0, src, bytes);
            if (!rc && bytes) memcpy(S.h.w_mvfd.data(), src, bytes);
            return rc;
        }
        case SPH_F_WALLS_ACC: return h2d(c, A.w_acc + w0, src, bytes);
        case SPH_F_BONDS_DELTA: return h2d(c, A.bonds_delta + bb, src, bytes);
        case SPH_F_BONDS_OMEGA: return h2d(c, A.bonds_omega + bb, src, bytes);
        case SPH_F_COLOR:
            if (bytes) memcpy(S.h.color.data(), src, bytes);
            return SPH_OK;
        default: return fail(c, SPH_E_ARG, "field %d is read-only", f);
    }
}

int sph_set_field(sph_ctx* c, int f, const void* src, size_t bytes) {
    return sph_set_field_sim(c, 0, f, src, bytes);
}

int sph_get_scalar_sim(sph_ctx* c, int sim, int which, double* v) {
    CHECK_SIM(c, sim);
    if (!v) return SPH_E_ARG;
    const Sim& S = c->sims[sim];
    switch (which) {
        case SPH_S_TIME:
        case SPH_S_ITER:
        case SPH_S_SORTS:
        case SPH_S_VL_BUILDS: {
            DevLoop L;
            int rc = read_loop(c, sim, &L);
            if (rc) return rc;
            *v = which == SPH_S_TIME    ? L.t
                 : which == SPH_S_ITER  ? (double)L.iter
                 : which == SPH_S_SORTS ? (double)L.sorts
                                        : (double)L.vl_builds;
            return SPH_OK;
        }
        case SPH_S_TIME_TOTAL: *v = S.h.t_total; return SPH_OK;
        case SPH_S_DT: *v = S.h.dt; return SPH_OK;
        case SPH_S_FILE_DT: *v = S.h.file_dt; return SPH_OK;
        case SPH_S_STEP_COUNT: *v = S.h.step_count; return SPH_OK;
        case SPH_S_NP: *v = S.h.np; return SPH_OK;
        case SPH_S_NW: *v = S.h.nw; return SPH_OK;
        case SPH_S_NB0: *v = S.h.nb0; return SPH_OK;
        case SPH_S_NCELLS: *v = S.P.ncells; return SPH_OK;
        case SPH_S_VL_SKIN: *v = S.P.vl_skin; return SPH_OK;
        case SPH_S_VL_CAP: *v = S.P.vl_cap; return SPH_OK;
        case SPH_S_CONTACT_LANES: *v = c->contact_lanes; return SPH_OK;
        case SPH_S_CONTACT_WG: *v = c->contact_wg; return SPH_OK;
        case SPH_S_RADIX_BITS: *v = c->radix_bits; return SPH_OK;
        case SPH_S_SORT_MODE: *v = c->sort_mode; return SPH_OK;
        case SPH_S_RESORT_MERGE_MAX: *v = S.P.mv_max; return SPH_OK;
        case SPH_S_NSIMS: *v = (double)c->sims.size(); return SPH_OK;
        case SPH_S_GRAPH_STEPS: *v = c->graph_steps; return SPH_OK;
        case SPH_S_XCD_ORDER: *v = c->A.xcd_order; return SPH_OK;
        case SPH_S_ADJ_CHECKPOINT: *v = (double)c->adj_checkpoint; return SPH_OK;
        case SPH_S_FUSED: *v = c->fused ? 1.0 : 0.0; return SPH_OK;
        case SPH_S_WALL_SUMS: *v = S.P.wall_sums; return SPH_OK;
        case SPH_S_K_N: *v = S.P.k_n; return SPH_OK;
        case SPH_S_BOND_K: *v = S.P.bond_k; return SPH_OK;
        case SPH_S_BOND_BREAK_EXT: *v = S.P.bond_break_ext; return SPH_OK;
        case SPH_S_INTEGRATE_LANES: *v = c->integrate_lanes; return SPH_OK;
        default: return fail(c, SPH_E_ARG, "unknown scalar %d", which);
    }
}

int sph_get_scalar(sph_ctx* c, int which, double* v) { return sph_get_scalar_sim(c, 0, which, v); }

int sph_set_scalar_sim(sph_ctx* c, int sim, int which, double v) {
    CHECK_SIM(c, sim);
    Sim& S = c->sims[sim];
    switch (which) {
        case SPH_S_TIME: {
            S.h.t = v;
            return h2d(c, &c->A.loop[sim].t, &v, 8);
        }
        case SPH_S_TIME_TOTAL: S.h.t_total = v; return SPH_OK;
        case SPH_S_DT:
            S.h.dt = v;
            S.P.dt = v;
            return upload_params(c);
        case SPH_S_FILE_DT: S.h.file_dt = v; return SPH_OK;
        case SPH_S_ITER: {
            // The per-step flags are kept by iteration parity, which all
            // simulations of a context share: the counter is set for all of
            // them, and the next step hashes again.
            const int64_t it = (int64_t)v;
            c->hash_pending = true;
            // The sorted copies the next step reads follow the counter's
            // parity: move them along when it changes.
            if (((it ^ c->sims[0].iter) & 1) != 0) {
                const size_t bytes = (size_t)c->A.ntot * 32;
                const int from = (int)(c->sims[0].iter & 1), to = from ^ 1;
                double4* const* sets[3] = {c->A.xs, c->A.vs, c->A.ws};
                for (auto set : sets)
                    HIPCHK(c, hipMemcpyAsync(set[to], set[from], bytes, hipMemcpyDeviceToDevice,
                                             c->stream), "copy");
            }
            for (size_t k = 0; k < c->sims.size(); ++k) {
                c->sims[k].iter = it;
                int rc = h2d(c, &c->A.loop[k].iter, &it, 8);
                if (rc) return rc;
            }
            return SPH_OK;
        }
        case SPH_S_STEP_COUNT: S.h.step_count = (uint32_t)v; return SPH_OK;
        case SPH_S_VL_SKIN: {
            set_skin(S.P, v);
            int rc = upload_params(c);
            if (rc) return rc;
            return set_loop_flag(c, sim, &DevLoop::force_rebuild);
        }
        case SPH_S_VL_CAP: {
            if (!(v >= 8.0 && v <= (double)NV) || v != (double)(int)v)
                return fail(c, SPH_E_ARG, "list capacity must be a whole number in [8, %d] "
                            "(got %g)", NV, v);
            S.P.vl_cap = (uint32_t)v;
            int rc = upload_params(c);
            if (rc) return rc;
            return set_loop_flag(c, sim, &DevLoop::force_rebuild);
        }
        case SPH_S_CONTACT_WG:
            if (v != 128.0 && v != 256.0 && v != 512.0)
                return fail(c, SPH_E_ARG, "contact workgroup size must be 128, 256 or 512 "
                            "(got %g)", v);
            if (c->contact_wg != (int)v) drop_graphs(c);
            c->contact_wg = (int)v;
            return SPH_OK;
        case SPH_S_RADIX_BITS:
            if (v != 4.0 && v != 6.0 && v != 8.0)
                return fail(c, SPH_E_ARG, "radix digit width must be 4, 6 or 8 bits (got %g)", v);
            if (c->radix_bits != (int)v) drop_graphs(c);
            c->radix_bits = (int)v;
            c->passes = (c->sorted_bits + c->radix_bits - 1) / c->radix_bits;
            return SPH_OK;
        case SPH_S_GRAPH_STEPS:
            if (!(v >= 0.0 && v <= 4096.0))
                return fail(c, SPH_E_ARG, "graph steps must be in [0, 4096] (got %g)", v);
            c->graph_steps = (int)v;
            return SPH_OK;
        case SPH_S_WALL_SUMS:
            if (v != 0.0 && v != 1.0)
                return fail(c, SPH_E_ARG, "wall sums must be 0 or 1 (got %g)", v);
            S.P.wall_sums = (uint32_t)v;
            return upload_params(c);
        case SPH_S_BOND_K:
            if (!(v >= 0.0 && v < 1e300))
                return fail(c, SPH_E_ARG, "bond stiffness must be finite and >= 0 (got %g)", v);
            S.P.bond_k = v;
            return upload_params(c);
        case SPH_S_BOND_BREAK_EXT:
            if (!(v >= 0.0 && v < 1e300))
                return fail(c, SPH_E_ARG, "bond break extension must be finite and >= 0 "
                            "(got %g)", v);
            S.P.bond_break_ext = v;
            return upload_params(c);
        case SPH_S_INTEGRATE_LANES:
            if (v != 1.0 && v != 4.0)
                return fail(c, SPH_E_ARG, "integrate lanes must be 1 or 4 (got %g)", v);
            if (c->integrate_lanes != (int)v) drop_graphs(c);
            c->integrate_lanes = (int)v;
            return SPH_OK;
        case SPH_S_FUSED:
            if (v != 0.0 && v != 1.0)
                return fail(c, SPH_E_ARG, "fused must be 0 or 1 (got %g)", v);
            drop_graphs(c);
            c->fused = v != 0.0;
            return SPH_OK;
        case SPH_S_ADJ_CHECKPOINT:
            if (!(v >= 0.0 && v <= 1e9) || v != (double)(uint64_t)v)
                return fail(c, SPH_E_ARG, "checkpoint interval must be a whole number >= 0 "
                            "(got %g)", v);
            c->adj_checkpoint = (uint64_t)v;
            return SPH_OK;
        case SPH_S_XCD_ORDER:
            if (v != 0.0 && v != 1.0)
                return fail(c, SPH_E_ARG, "XCD order must be 0 or 1 (got %g)", v);
            drop_graphs(c);
            c->A.xcd_order = (uint32_t)v;
            return SPH_OK;
        case SPH_S_CONTACT_LANES:
            if (v != 1.0 && v != 4.0 && v != 8.0 && v != 16.0)
                return fail(c, SPH_E_ARG, "contact lanes must be 1, 4, 8 or 16 (got %g)", v);
            if (c->contact_lanes != (int)v) drop_graphs(c);  // captured with the old kernel
            c->contact_lanes = (int)v;
            return SPH_OK;
        case SPH_S_SORT_MODE:
            if (v != 0.0 && v != 1.0 && v != 2.0)
                return fail(c, SPH_E_ARG, "sort mode must be 0, 1 or 2 (got %g)", v);
            if (c->sort_mode != (int)v) drop_graphs(c);
            c->sort_mode = (int)v;
            if (c->sort_mode != 2) c->resort_at_end = c->sort_mode == 1;
            return SPH_OK;
        case SPH_S_RESORT_MERGE_MAX:
            if (!(v >= 0.0 && v <= (double)MV_CAP))
                return fail(c, SPH_E_ARG, "merge limit must be in [0, %u] (got %g)", MV_CAP, v);
            S.P.mv_max = (uint32_t)v;
            return upload_params(c);
        default: return fail(c, SPH_E_ARG, "scalar %d is read-only", which);
    }
}

int sph_set_scalar(sph_ctx* c, int which, double v) { return sph_set_scalar_sim(c, 0, which, v); }

int sph_get_contacts(sph_ctx* c, uint32_t* idx, double* delta_t, double* distmod) {
    if (!c) return SPH_E_ARG;
    const size_t np = c->h0().np;
    int rc = SPH_OK;
    if (idx) rc = sph_get_field(c, SPH_F_CONTACTS, idx, np * NC * 4);
    if (!rc && delta_t) rc = sph_get_field(c, SPH_F_DELTA_T, delta_t, np * NC * 24);
    if (!rc && distmod) rc = sph_get_field(c, SPH_F_DISTMOD, distmod, np * NC * 32);
    return rc;
}

int sph_get_cells(sph_ctx* c, uint32_t* cell_id, uint32_t* perm, uint32_t* cell_start,
                  uint32_t* cell_end) {
    if (!c) return SPH_E_ARG;
    const size_t np = c->h0().np, nc = c->sims[0].P.ncells;
    int rc = SPH_OK;
    if (cell_id) rc = sph_get_field(c, SPH_F_CELLID, cell_id, np * 4);
    if (!rc && perm) rc = sph_get_field(c, SPH_F_PERM, perm, np * 4);
    if (!rc && cell_start) rc = sph_get_field(c, SPH_F_CELLSTART, cell_start, nc * 4);
    if (!rc && cell_end) rc = sph_get_field(c, SPH_F_CELLEND, cell_end, nc * 4);
    return rc;
}

// The adjoint of nsteps steps, over every simulation of the context.  Host
// arrays concatenate the simulations' particles (rows hbase .. hbase + np
// of simulation s).  Seeds: dL_dx/dv/dw on the final state, and with
// every > 0 the (nsteps / every) x rows x 3 arrays sx / sv on the positions
// / velocities after every, 2 every, ... steps (any may be null).
// g_params receives nsims x ADJ_NPARAM values.
// phase 0 runs the forward and the reverse pass; phase 1 only the forward
// pass (recording, or saving checkpoints), with the positions after every,
// 2 every, ... steps into samples_x (nsamples x rows x 3, may be null), and
// keeps what the reverse pass needs in c->adj_plan; phase 2 only the
// reverse pass of the run phase 1 recorded, which must be the context's
// last operation.
static int adjoint_run(sph_ctx* c, uint64_t nsteps, const double* dL_dx, const double* dL_dv,
                       const double* dL_dw, uint64_t every, const double* sx,
                       const double* sv, double* g_x0, double* g_v0, double* g_w0,
                       double* g_params, int phase = 0, double* samples_x = nullptr) {
    if (!c) return SPH_E_ARG;
    AdjPlan* plan = nullptr;
    if (phase == 2) {
        plan = c->adj_plan;
        if (!plan || plan->nsteps != nsteps || plan->every != every ||
            plan->iter_end != c->sims[0].iter || plan->buf != c->adj_buf)
            return fail(c, SPH_E_ARG, "no recorded forward pass of this run to reverse (the "
                        "context must not step or record again in between)");
    } else {
        delete c->adj_plan;
        c->adj_plan = nullptr;
    }
    const size_t nsim = c->sims.size();
    size_t rows = 0, nwt = 0;
    for (const Sim& S : c->sims) {
        const Host& h = S.h;
        if (h.contactmodel != 2 || h.kappa != 0.0 || h.E > 0.001)
            return fail(c, SPH_E_UNSUPPORTED,
                        "the adjoint covers contact model 2 without cohesion or E-scaled "
                        "stiffness");
        for (uint32_t w = 0; w < h.nw; ++w)
            if (h.wmode[w] < 0 || h.wmode[w] > 2)
                return fail(c, SPH_E_UNSUPPORTED,
                            "the adjoint covers fixed, stress-controlled and velocity-controlled "
                            "walls (wmode 0, 1, 2), not wmode %d", h.wmode[w]);
        for (uint32_t i = 0; i < h.np; ++i)
            if (h.vel[4 * i + 3] > 5.0 && h.vel[4 * i + 3] < 10.0)
                return fail(c, SPH_E_UNSUPPORTED, "the adjoint excludes rate-state particles");
        rows += h.np;
        nwt += h.nw;
    }
    if (c->shear_stress || c->any_adaptive)
        return fail(c, SPH_E_UNSUPPORTED,
                    "the adjoint excludes a wmode-3 wall and an adaptive grid");
    if (nsteps == 0) return SPH_OK;
    DevArrays& A = c->A;
    const size_t ntot = A.ntot, nc = ntot * NC, nb = A.nbond;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    // Tape per step: x, v, w at the step start and the order (sorted), the
    // displacements of the step's contacts before it and their partners'
    // sorted indices (written by the contact pass, at their slots), the
    // walls, the rows the contact pass left, and the bonds' inputs.
    const size_t per_step = up(ntot * 3 * sizeof(double4)) + up(ntot * 4) +
                            up(nc * sizeof(double4)) + up(nc * 4) +
                            up(2 * nwt * sizeof(double4)) + up(nc * 4) +
                            (nb ? up(nb * 8 * sizeof(double4)) + up(nb * sizeof(uint4)) +
                                      up(nsim * 4)
                                : 0);
    // Reverse state: cotangents of x, v, w (two sets), of the two
    // accelerations, of the displacements (two sets), of the bond state
    // (two sets), the bonds' contributions and parameter shares, the
    // workgroup partials, the parameter sums, the walls, the staging rows.
    const size_t nwg = ntot / ADJ_PPB;
    const size_t state_b = up(ntot * 10 * sizeof(double4)) + up(2 * nc * sizeof(double4)) +
                           up(4 * nb * sizeof(double4)) + up(6 * nb * sizeof(double4)) +
                           up(nb * ADJ_NBPARAM * 8) + up(nwg * ADJ_NPART * 8) +
                           up(nsim * ADJ_NPARAM * 8) + up(nwt * ADJ_WALL * 8 + 8) +
                           up(rows * 3 * sizeof(double)) + up(nc * 4) + 256;
    // A checkpoint: every device allocation of the context.
    size_t snap_b = 0;
    for (size_t b : c->alloc_bytes) snap_b += up(b);
    // Segment length K: the whole run, unless set or unless the whole run's
    // tape does not fit; then the longest segments that fit.
    size_t free_b = 0, total_b = 0;
    HIPCHK(c, hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const double avail = 0.9 * (double)(free_b + c->adj_bytes);
    auto need_for = [&](uint64_t K) {
        const uint64_t nseg = (nsteps + K - 1) / K;
        return (double)per_step * (double)K + (double)state_b +
               (nseg > 1 ? (double)(nseg + 1) * (double)snap_b : 0.0);
    };
    uint64_t K = c->adj_checkpoint > 0 ? std::min<uint64_t>(c->adj_checkpoint, nsteps) : nsteps;
    if (plan) K = plan->K;
    else if (c->adj_checkpoint == 0 && need_for(K) > avail) {
        K = 0;
        for (uint64_t k = nsteps; k >= 1; k = k > 1 ? k * 3 / 4 : 0) {
            if (need_for(k) <= avail) {
                K = k;
                break;
            }
            if (k == 1) break;
        }
        if (K == 0)
            return fail(c, SPH_E_ARG, "the adjoint of %llu steps does not fit in device memory",
                        (unsigned long long)nsteps);
    }
    const uint64_t nseg = (nsteps + K - 1) / K;
    const double need = need_for(K);
    if (!plan && need > avail)
        return fail(c, SPH_E_ARG, "recording %llu steps in segments of %llu needs %.1f GB of "
                    "device memory", (unsigned long long)nsteps, (unsigned long long)K,
                    need / 1e9);
    if (need > (double)c->adj_bytes) {
        HIPCHK(c, hipStreamSynchronize(c->stream), "sync");
        if (c->adj_buf) (void)hipFree(c->adj_buf);
        c->adj_buf = nullptr;
        c->adj_bytes = 0;
        if (hipMalloc(&c->adj_buf, (size_t)need) != hipSuccess) {
            (void)hipGetLastError();
            c->adj_buf = nullptr;
            return fail(c, SPH_E_DEVICE, "hipMalloc failed for the adjoint tape (%.1f GB)",
                        need / 1e9);
        }
        c->adj_bytes = (size_t)need;
    }
    hipStream_t s = c->stream;
    char* tape = static_cast<char*>(c->adj_buf);
    struct Rec {
        double4 *x, *v, *w, *dt0, *wnx, *wmv, *bond_in;
        uint32_t *perm, *srt, *rows, *bond_full;
        uint4* bond_idx;
    };
    std::vector<Rec> rec(K);
    for (uint64_t k = 0; k < K; ++k) {
        char* q = tape + k * per_step;
        Rec& r = rec[k];
        r.x = reinterpret_cast<double4*>(q);
        r.v = r.x + ntot;
        r.w = r.v + ntot;
        q += up(ntot * 3 * sizeof(double4));
        r.perm = reinterpret_cast<uint32_t*>(q);
        q += up(ntot * 4);
        r.dt0 = reinterpret_cast<double4*>(q);
        q += up(nc * sizeof(double4));
        r.srt = reinterpret_cast<uint32_t*>(q);
        q += up(nc * 4);
        r.wnx = reinterpret_cast<double4*>(q);
        r.wmv = r.wnx + nwt;
        q += up(2 * nwt * sizeof(double4));
        r.rows = reinterpret_cast<uint32_t*>(q);
        q += up(nc * 4);
        r.bond_in = nullptr;
        r.bond_idx = nullptr;
        r.bond_full = nullptr;
        if (nb) {
            r.bond_in = reinterpret_cast<double4*>(q);
            q += up(nb * 8 * sizeof(double4));
            r.bond_idx = reinterpret_cast<uint4*>(q);
            q += up(nb * sizeof(uint4));
            r.bond_full = reinterpret_cast<uint32_t*>(q);
        }
    }
    const int64_t iter0 = plan ? plan->iter0 : c->sims[0].iter;
    // Reverse-pass state, and the seeds (uploaded before the forward pass,
    // so that only the gradients' download waits for the device).
    char* q = tape + K * per_step;
    auto take = [&](size_t bytes) {
        char* r = q;
        q += up(bytes);
        return r;
    };
    double4* L[2][3];
    double4 *LA[2], *LAA[2];
    {
        double4* base = reinterpret_cast<double4*>(take(ntot * 10 * sizeof(double4)));
        for (int b = 0; b < 2; ++b) {
            for (int f = 0; f < 3; ++f) {
                L[b][f] = base;
                base += ntot;
            }
            LA[b] = base;
            base += ntot;
            LAA[b] = base;
            base += ntot;
        }
    }
    AdjState g;
    double4* ldt0 = reinterpret_cast<double4*>(take(2 * nc * sizeof(double4)));
    double4* ldt[2] = {ldt0, ldt0 + nc};
    double4* lb0 = reinterpret_cast<double4*>(take(4 * nb * sizeof(double4)));
    double4* lb[2] = {lb0, lb0 + 2 * nb};
    g.bcontrib = reinterpret_cast<double4*>(take(6 * nb * sizeof(double4)));
    g.bpart = reinterpret_cast<double*>(take(nb * ADJ_NBPARAM * 8));
    g.part = reinterpret_cast<double*>(take(nwg * ADJ_NPART * 8));
    g.gparam = reinterpret_cast<double*>(take(nsim * ADJ_NPARAM * 8));
    g.wall = reinterpret_cast<double*>(take(nwt * ADJ_WALL * 8 + 8));
    g.asym = reinterpret_cast<uint32_t*>(g.wall + nwt * ADJ_WALL);
    double* stage = reinterpret_cast<double*>(take(rows * 3 * sizeof(double)));
    // A checkpoint's rows with the partners' original indices (the tape's).
    uint32_t* prev_orig = reinterpret_cast<uint32_t*>(take(nc * 4));
    char* snaps = tape + K * per_step + state_b;  // (nseg + 1) checkpoints when nseg > 1
    AdjStep base_st;
    memset(&base_st, 0, sizeof base_st);
    base_st.ptab = c->d_ptab;
    base_st.chunk_sim = A.chunk_sim;
    base_st.nsim = (uint32_t)nsim;
    base_st.ntot = (uint32_t)ntot;
    base_st.nbond = (uint32_t)nb;
    base_st.bond_sim = A.bond_sim;
    base_st.badj_off = A.badj_off;
    base_st.badj = A.badj;
    const double* seeds[3] = {dL_dx, dL_dv, dL_dw};
    int rc = SPH_OK;
    hipError_t e = hipSuccess;
    for (int f = 0; f < 3 && !rc && phase != 1; ++f) {
        if (!seeds[f]) {
            e = hipMemsetAsync(L[0][f], 0, ntot * 32, s);
            if (e != hipSuccess) return hip_fail(c, e, "memset");
            continue;
        }
        if (rows) rc = h2d(c, stage, seeds[f], rows * 24);
        if (!rc) {
            e = launch_adj_expand(L[0][f], stage, base_st, s);
            if (e != hipSuccess) return hip_fail(c, e, "seed");
        }
    }
    if (rc) return rc;
    if (phase != 1) {
        e = hipMemsetAsync(LA[0], 0, 2 * ntot * 32, s);  // LA[0], LAA[0]
        if (e == hipSuccess) e = hipMemsetAsync(g.gparam, 0, nsim * ADJ_NPARAM * 8, s);
        if (e == hipSuccess) e = hipMemsetAsync(g.wall, 0, nwt * ADJ_WALL * 8 + 8, s);
        if (e == hipSuccess && nb) e = hipMemsetAsync(lb0, 0, 4 * nb * sizeof(double4), s);
        if (e != hipSuccess) return hip_fail(c, e, "memset");
    }

    // Checkpoints: the device allocations, and the host state enqueue_step
    // reads (the loop mirrors and the hash/re-sort bookkeeping).
    std::vector<AdjHostMark> marks;
    if (plan) marks = plan->marks;
    else marks.resize(nseg > 1 ? nseg + 1 : 0);
    using HostMark = AdjHostMark;
    size_t crow_off = 0, prow_off = 0;
    for (size_t k = 0, off = 0; k < c->allocs.size(); off += up(c->alloc_bytes[k]), ++k) {
        if (c->allocs[k] == (void*)A.crow) crow_off = off;
        if (c->allocs[k] == (void*)A.perm_rows) prow_off = off;
    }
    auto checkpoint = [&](uint64_t m, bool save) -> int {
        char* cbase = snaps + m * snap_b;
        size_t off = 0;
        for (size_t k = 0; k < c->allocs.size(); ++k) {
            void* dev = c->allocs[k];
            hipError_t e_ = save ? hipMemcpyAsync(cbase + off, dev, c->alloc_bytes[k],
                                                  hipMemcpyDeviceToDevice, s)
                                 : hipMemcpyAsync(dev, cbase + off, c->alloc_bytes[k],
                                                  hipMemcpyDeviceToDevice, s);
            if (e_ != hipSuccess) return hip_fail(c, e_, "checkpoint");
            off += up(c->alloc_bytes[k]);
        }
        HostMark& hm = marks[m];
        if (save) {
            hm.t.resize(nsim);
            hm.iter.resize(nsim);
            for (size_t k = 0; k < nsim; ++k) {
                hm.t[k] = c->sims[k].h.t;
                hm.iter[k] = c->sims[k].iter;
            }
            hm.hash_pending = c->hash_pending;
            hm.prev_walls_resorted = c->prev_walls_resorted;
        } else {
            for (size_t k = 0; k < nsim; ++k) {
                c->sims[k].h.t = hm.t[k];
                c->sims[k].iter = hm.iter[k];
            }
            c->hash_pending = hm.hash_pending;
            c->prev_walls_resorted = hm.prev_walls_resorted;
        }
        return SPH_OK;
    };
    // Forward steps [k0, k0 + n) with the tape; the contact pass of each step
    // records the state and walls the step starts from, its rows and the
    // displacements it read, the bond pass the bonds' inputs.
    // Positions after global step gk (phase 1 with samples_x): sample m =
    // (gk + 1) / every - 1, the simulations' particles concatenated.
    std::vector<double> xtmp;
    auto sample = [&](uint64_t gk) -> int {
        if (phase != 1 || !samples_x || every == 0 || (gk + 1) % every != 0) return SPH_OK;
        double* dst = samples_x + ((gk + 1) / every - 1) * rows * 3;
        for (const Sim& S : c->sims) {
            const size_t np = S.h.np;
            xtmp.resize(np * 4);
            const int rc_ = np ? d2h(c, xtmp.data(), A.x + S.P.base, np * 32) : SPH_OK;
            if (rc_) return rc_;
            for (size_t i = 0; i < np; ++i)
                for (int d = 0; d < 3; ++d) dst[(S.P.hbase + i) * 3 + d] = xtmp[4 * i + d];
        }
        return SPH_OK;
    };
    auto record = [&](uint64_t n, uint64_t k0) -> int {
        int rc_ = SPH_OK;
        for (uint64_t k = 0; k < n && !rc_; ++k) {
            const Rec& r = rec[k];
            A.tape_rows = r.rows;
            A.tape_dt0 = r.dt0;
            A.tape_srt = r.srt;
            A.tape_perm = r.perm;
            A.tape_x = r.x;
            A.tape_v = r.v;
            A.tape_w = r.w;
            A.tape_wnx = r.wnx;
            A.tape_wmv = r.wmv;
            A.tape_bond_in = r.bond_in;
            A.tape_bond_idx = r.bond_idx;
            A.tape_bond_full = r.bond_full;
            rc_ = enqueue_step(c);
            A.tape_rows = A.tape_srt = A.tape_perm = A.tape_bond_full = nullptr;
            A.tape_dt0 = A.tape_x = A.tape_v = A.tape_w = A.tape_wnx = A.tape_wmv = nullptr;
            A.tape_bond_in = nullptr;
            A.tape_bond_idx = nullptr;
            if (!rc_) rc_ = sample(k0 + k);
        }
        return rc_;
    };
    // Reverse of the recorded steps [k0, k0 + n), last first.  prev0: the
    // rows before step k0 (null for the run's first step).
    int cur = 0;
    auto reverse = [&](uint64_t k0, uint64_t n, const uint32_t* prev0) -> int {
        for (uint64_t kk = n; kk-- > 0;) {
            const Rec& r = rec[kk];
            const uint64_t gk = k0 + kk;
            AdjStep st = base_st;
            st.x = r.x;
            st.v = r.v;
            st.w = r.w;
            st.perm = r.perm;
            st.dt0 = r.dt0;
            st.srt = r.srt;
            st.rows = r.rows;
            st.prev_rows = kk > 0 ? rec[kk - 1].rows : prev0;
            st.wall_nx = r.wnx;
            st.wall_mvfd = r.wmv;
            st.bond_in = r.bond_in;
            st.bond_idx = r.bond_idx;
            st.bond_full = r.bond_full;
            st.first = (iter0 + (int64_t)gk == 0) ? 1 : 0;
            st.last = gk + 1 == nsteps ? 1 : 0;
            // A loss term on the state after this step: added to its
            // cotangent before the step is reversed.
            if (every > 0 && (gk + 1) % every == 0) {
                const size_t m = (size_t)((gk + 1) / every - 1);
                const double* samp[2] = {sx, sv};
                for (int f = 0; f < 2; ++f) {
                    if (!samp[f] || !rows) continue;
                    const int rc_ = h2d(c, stage, samp[f] + m * rows * 3, rows * 24);
                    if (rc_) return rc_;
                    hipError_t ea = launch_adj_accumulate(L[cur][f], stage, st, s);
                    if (ea != hipSuccess) return hip_fail(c, ea, "sample seed");
                }
            }
            for (int f = 0; f < 3; ++f) {
                g.Lout[f] = L[cur][f];
                g.Lin[f] = L[cur ^ 1][f];
            }
            g.La_out = LA[cur];
            g.Laa_out = LAA[cur];
            g.La_in = LA[cur ^ 1];
            g.Laa_in = LAA[cur ^ 1];
            g.Ldt_in = ldt[cur];
            g.Ldt_out = ldt[cur ^ 1];
            g.Lb_out = lb[cur];
            g.Lb_in = lb[cur ^ 1];
            hipError_t e_ = launch_adj_step(st, g, s);
            if (e_ != hipSuccess) return hip_fail(c, e_, "adjoint step");
            cur ^= 1;
        }
        return SPH_OK;
    };
    // The forward pass: the whole tape, or checkpoints (m holds the state
    // before step m K; m = nseg the end).
    if (phase != 2) {
        if (nseg == 1) {
            rc = record(nsteps, 0);
        } else {
            for (uint64_t m = 0; m < nseg && !rc; ++m) {
                rc = checkpoint(m, true);
                const uint64_t n = std::min<uint64_t>(K, nsteps - m * K);
                for (uint64_t k = 0; k < n && !rc; ++k) {
                    rc = enqueue_step(c);
                    if (!rc) rc = sample(m * K + k);
                }
            }
            if (!rc) rc = checkpoint(nseg, true);
        }
    }
    if (phase == 1) {
        if (rc) return rc;
        HIPCHK(c, hipStreamSynchronize(s), "sync");
        AdjPlan* pl = new AdjPlan();
        pl->nsteps = nsteps;
        pl->every = every;
        pl->K = K;
        pl->iter0 = iter0;
        pl->iter_end = c->sims[0].iter;
        pl->buf = c->adj_buf;
        pl->marks = std::move(marks);
        c->adj_plan = pl;
        return SPH_OK;
    }
    if (phase == 2) {
        delete c->adj_plan;
        c->adj_plan = nullptr;
    }
    // The reverse pass.
    if (nseg == 1) {
        if (!rc) rc = reverse(0, nsteps, nullptr);
    } else {
        for (uint64_t m = nseg; m-- > 0 && !rc;) {
            const uint64_t n = std::min<uint64_t>(K, nsteps - m * K);
            rc = checkpoint(m, false);
            if (!rc) rc = record(n, m * K);
            // The rows before step m K: those the checkpoint's contact rows
            // hold, with original indices.
            const uint32_t* prev0 = nullptr;
            if (m > 0 && !rc) {
                const char* ck = snaps + m * snap_b;
                e = launch_rows_to_orig(reinterpret_cast<const uint32_t*>(ck + crow_off),
                                        reinterpret_cast<const uint32_t*>(ck + prow_off),
                                        prev_orig, (uint32_t)nc, s);
                if (e != hipSuccess) return hip_fail(c, e, "checkpoint rows");
                prev0 = prev_orig;
            }
            if (!rc) rc = reverse(m * K, n, prev0);
        }
        if (!rc) rc = checkpoint(nseg, false);
    }
    if (rc) return rc;
    double* outs[3] = {g_x0, g_v0, g_w0};
    for (int f = 0; f < 3 && !rc; ++f) {
        if (!outs[f] || !rows) continue;
        e = launch_adj_pack(stage, L[cur][f], base_st, s);
        if (e != hipSuccess) return hip_fail(c, e, "pack");
        rc = d2h(c, outs[f], stage, rows * 24);
    }
    uint32_t asym = 0;
    if (!rc) rc = d2h(c, &asym, g.asym, 4);
    if (!rc && asym)
        return fail(c, SPH_E_UNSUPPORTED,
                    "%u recorded contacts were missing from their partner's row (a contact row "
                    "overflowed NC = %d); the gradients are incomplete",
                    asym, NC);
    if (!rc && g_params) rc = d2h(c, g_params, g.gparam, nsim * ADJ_NPARAM * 8);
    return rc;
}

static int single_sim_adjoint(sph_ctx* c) {
    if (!c) return SPH_E_ARG;
    if (c->sims.size() != 1)
        return fail(c, SPH_E_UNSUPPORTED,
                    "sph_adjoint / sph_adjoint_samples act on a single-simulation context; "
                    "use sph_adjoint_batch / sph_adjoint_samples_batch for a batch");
    return SPH_OK;
}

static int check_samples(sph_ctx* c, uint64_t every, uint64_t nsamples) {
    if (every == 0 || nsamples == 0 || nsamples > (uint64_t)1e12 / every)
        return fail(c, SPH_E_ARG, "sampling every %llu steps, %llu times: both must be "
                    "positive", (unsigned long long)every, (unsigned long long)nsamples);
    return SPH_OK;
}

int sph_adjoint(sph_ctx* c, uint64_t nsteps, const double* dL_dx, const double* dL_dv,
                const double* dL_dw, double* g_x0, double* g_v0, double* g_w0,
                double* g_params) {
    const int rc = single_sim_adjoint(c);
    if (rc) return rc;
    return adjoint_run(c, nsteps, dL_dx, dL_dv, dL_dw, 0, nullptr, nullptr, g_x0, g_v0, g_w0,
                       g_params);
}

int sph_adjoint_samples(sph_ctx* c, uint64_t every, uint64_t nsamples, const double* dL_dx,
                        const double* dL_dv, double* g_x0, double* g_v0, double* g_w0,
                        double* g_params) {
    int rc = single_sim_adjoint(c);
    if (!rc) rc = check_samples(c, every, nsamples);
    if (rc) return rc;
    return adjoint_run(c, every * nsamples, nullptr, nullptr, nullptr, every, dL_dx, dL_dv,
                       g_x0, g_v0, g_w0, g_params);
}

int sph_adjoint_record(sph_ctx* c, uint64_t every, uint64_t nsamples, double* samples_x) {
    if (!c) return SPH_E_ARG;
    const int rc = check_samples(c, every, nsamples);
    if (rc) return rc;
    return adjoint_run(c, every * nsamples, nullptr, nullptr, nullptr, every, nullptr, nullptr,
                       nullptr, nullptr, nullptr, nullptr, 1, samples_x);
}

int sph_adjoint_reverse(sph_ctx* c, uint64_t every, uint64_t nsamples, const double* dL_dx,
                        const double* dL_dv, double* g_x0, double* g_v0, double* g_w0,
                        double* g_params) {
    if (!c) return SPH_E_ARG;
    const int rc = check_samples(c, every, nsamples);
    if (rc) return rc;
    return adjoint_run(c, every * nsamples, nullptr, nullptr, nullptr, every, dL_dx, dL_dv,
                       g_x0, g_v0, g_w0, g_params, 2, nullptr);
}

int sph_adjoint_batch(sph_ctx* c, uint64_t nsteps, const double* dL_dx, const double* dL_dv,
                      const double* dL_dw, double* g_x0, double* g_v0, double* g_w0,
                      double* g_params) {
    if (!c) return SPH_E_ARG;
    return adjoint_run(c, nsteps, dL_dx, dL_dv, dL_dw, 0, nullptr, nullptr, g_x0, g_v0, g_w0,
                       g_params);
}

int sph_adjoint_samples_batch(sph_ctx* c, uint64_t every, uint64_t nsamples,
                              const double* dL_dx, const double* dL_dv, double* g_x0,
                              double* g_v0, double* g_w0, double* g_params) {
    if (!c) return SPH_E_ARG;
    const int rc = check_samples(c, every, nsamples);
    if (rc) return rc;
    return adjoint_run(c, every * nsamples, nullptr, nullptr, nullptr, every, dL_dx, dL_dv,
                       g_x0, g_v0, g_w0, g_params);
}

int sph_autotune(sph_ctx* c, uint64_t steps_per_trial, double* us_per_step) {
    if (!c || steps_per_trial == 0) return SPH_E_ARG;
    // Every setting tried gives the same results bit for bit, so the trials
    // are ordinary steps of the simulation; only the kernels' times differ.
    // The contact settings are judged by the contact kernel's mean time
    // (HIP events around each launch), which the occasional re-sort or list
    // rebuild of a trial does not disturb; the radix digit width by the
    // time of full sorts of the current keys (which reproduce the current
    // order).  A trial first runs half as many untimed steps (the list
    // rebuild a new capacity forces), and a setting replaces the incumbent
    // only when at least 3% faster.
    const uint64_t warm = std::max<uint64_t>(1, steps_per_trial / 2);
    const bool history = c->sims[0].h.contactmodel == 2 || c->sims[0].h.contactmodel == 3;
    auto contact_trial = [&](double* us) -> int {
        int rc = sph_step(c, warm);
        if (!rc) rc = sph_sync(c);
        if (rc) return rc;
        prof_collect(c);
        c->prof_names.clear();
        c->prof_ms.clear();
        c->prof_n.clear();
        const bool was = c->profile;
        c->profile = true;
        rc = sph_step(c, steps_per_trial);
        if (!rc) rc = sph_sync(c);
        prof_collect(c);
        c->profile = was;
        if (rc) return rc;
        *us = 0.0;
        for (size_t k = 0; k < c->prof_names.size(); ++k)
            if ((c->prof_names[k] == "contact" || c->prof_names[k] == "contact_integrate") &&
                c->prof_n[k] > 0)
                *us = 1000.0 * c->prof_ms[k] / (double)c->prof_n[k];
        return *us > 0.0 ? SPH_OK : fail(c, SPH_E_DEVICE, "autotune: no contact launches timed");
    };
    auto sort_trial = [&](double* us) -> int {
        // Full sorts of the keys of the next step: the same permutation and
        // cell ranges the step would use; the flag that lets the sort
        // kernels run is the one the next step consumes and retires.
        hipStream_t s = c->stream;
        const uint32_t slot = (uint32_t)(c->sims[0].iter & 1);
        const uint32_t one = 1u;
        hipEvent_t e0 = take_event(c), e1 = take_event(c);
        constexpr int REPS = 4;
        float total = 0.0f;
        for (int r = 0; r < REPS; ++r) {
            HIPCHK(c, hipMemcpyAsync(&c->A.glob->changed[slot], &one, 4, hipMemcpyHostToDevice,
                                     s), "flag");
            HIPCHK(c, hipEventRecord(e0, s), "event");
            HIPCHK(c, launch_sort(c->A, c->passes, c->radix_bits, s), "sort");
            HIPCHK(c, launch_cells(c->A, c->sorted_bits, s), "cells");
            HIPCHK(c, hipEventRecord(e1, s), "event");
            HIPCHK(c, hipEventSynchronize(e1), "sync");
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            total += ms;
        }
        prof_state(c).pool.push_back(e0);
        prof_state(c).pool.push_back(e1);
        *us = 1000.0 * total / REPS;
        return SPH_OK;
    };
    int rc = SPH_OK;
    // Coordinate search over one knob at a time: each candidate value is
    // tried from the incumbent settings.  (The candidate-list skin is not
    // tuned here: its cost is the rebuild rate, which short trials do not
    // see.)  The list capacity is set on every simulation of the context.
    auto set_all = [&](int which, double v) -> int {
        if (which == SPH_S_VL_CAP) {
            for (size_t k = 0; k < c->sims.size(); ++k) {
                const int r = sph_set_scalar_sim(c, (int)k, which, v);
                if (r) return r;
            }
            return SPH_OK;
        }
        return sph_set_scalar(c, which, v);
    };
    double best = 0.0;
    auto search = [&](int which, std::initializer_list<double> values, bool sort) {
        double cur = 0.0;
        if (rc || (rc = sph_get_scalar(c, which, &cur))) return;
        rc = sort ? sort_trial(&best) : contact_trial(&best);  // the incumbent
        for (double v : values) {
            if (v == cur || rc) continue;
            rc = set_all(which, v);
            double t = 0.0;
            if (!rc) rc = sort ? sort_trial(&t) : contact_trial(&t);
            if (!rc && t < 0.97 * best) {
                best = t;
                cur = v;
            } else if (!rc) {
                rc = set_all(which, cur);
            }
        }
    };
    if (history) {
        search(SPH_S_CONTACT_LANES, {4.0, 8.0, 16.0}, false);
        if (!rc && c->contact_lanes == 4) search(SPH_S_CONTACT_WG, {128.0, 256.0, 512.0}, false);
        search(SPH_S_VL_CAP, {16.0, 24.0, 32.0}, false);
    }
    search(SPH_S_XCD_ORDER, {0.0, 1.0}, false);
    double contact_best = best;
    search(SPH_S_RADIX_BITS, {4.0, 6.0, 8.0}, true);
    if (!rc && us_per_step) *us_per_step = contact_best;
    return rc;
}

int sph_profile(sph_ctx* c, int enable) {
    if (!c) return SPH_E_ARG;
    prof_collect(c);
    c->profile = enable != 0;
    c->prof_names.clear();
    c->prof_ms.clear();
    c->prof_n.clear();
    return SPH_OK;
}

int sph_profile_report(sph_ctx* c, char* buf, size_t cap) {
    if (!c) return SPH_E_ARG;
    prof_collect(c);
    std::string out;
    char line[256];
    for (size_t k = 0; k < c->prof_names.size(); ++k) {
        snprintf(line, sizeof line, "%s %.3f %ld\n", c->prof_names[k].c_str(),
                 c->prof_n[k] ? 1000.0 * c->prof_ms[k] / c->prof_n[k] : 0.0, c->prof_n[k]);
        out += line;
    }
    if (buf && cap) snprintf(buf, cap, "%s", out.c_str());
    return (int)out.size();
}

void sph_close(sph_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    if (c->sample_ev) (void)hipEventDestroy(c->sample_ev);
    if (c->loop_sample) (void)hipHostFree(c->loop_sample);
    if (c->prof) {
        prof_collect(c);
        for (hipEvent_t e : c->prof->pool) (void)hipEventDestroy(e);
        delete c->prof;
    }
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->adj_buf) (void)hipFree(c->adj_buf);
    delete c->adj_plan;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* sph_last_error(const sph_ctx* c) { return c ? c->err.c_str() : g_open_error.c_str(); }
const char* sph_open_error(void) { return g_open_error.c_str(); }
const char* sph_version_string(void) { return "sphere_hip 0.1 (.bin 2.15, gfx950)"; }

int sph_find_overlaps(sph_ctx* c, int sim, uint32_t* pairs, double* overlaps, uint64_t cap,
                      uint64_t* count) {
    CHECK_SIM(c, sim);
    if (!count || (cap && (!pairs || !overlaps))) return SPH_E_ARG;
    const Sim& S = c->sims[sim];
    const uint32_t np = S.h.np, b0 = S.P.base;
    const double4* x = c->A.x + b0;
    const double4* vel = c->A.vel + b0;
    hipStream_t s = c->stream;
    uint32_t* counts = nullptr;
    uint64_t* offs = nullptr;
    uint32_t* dpairs = nullptr;
    double* ddelta = nullptr;
    double* dbounds = nullptr;
    uint32_t* key = nullptr;
    uint32_t* cells = nullptr;
    uint32_t* cursor = nullptr;
    uint32_t* order = nullptr;
    auto release = [&]() {
        (void)hipStreamSynchronize(s);
        for (void* p : {(void*)counts, (void*)offs, (void*)dpairs, (void*)ddelta, (void*)dbounds,
                        (void*)key, (void*)cells, (void*)cursor, (void*)order})
            (void)hipFree(p);
    };
#define AFAIL(e_, what)                        \
    do {                                       \
        hipError_t ee_ = (e_);                 \
        if (ee_ != hipSuccess) {               \
            release();                         \
            return hip_fail(c, ee_, what);     \
        }                                      \
    } while (0)
    AFAIL(hipMalloc(&counts, (size_t)np * 4 + 4), "hipMalloc");
    AFAIL(hipMalloc(&offs, (size_t)np * 8 + 8), "hipMalloc");

    // A uniform grid whose cells are at least the largest contact distance
    // (2 r_max, with a margin for the rounding of the cell coordinates)
    // holds every overlapping pair within adjacent cells.  The grid is
    // coarsened until it has at most max(2 np, 4096) cells.  Positions that
    // are not finite fall back to the all-pairs search.
    double bounds[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (np >= 2) {
        AFAIL(hipMalloc(&dbounds, 64), "hipMalloc");
        AFAIL(launch_overlap_bounds(x, np, dbounds, s), "overlap bounds");
        AFAIL(hipMemcpyAsync(bounds, dbounds, 64, hipMemcpyDeviceToHost, s), "D2H");
        AFAIL(hipStreamSynchronize(s), "sync");
    }
    const bool use_grid = np >= 2 && bounds[7] == 1.0;
    OverlapGrid g{};
    if (use_grid) {
        double maxabs = 0.0, emax = 0.0, ext[3];
        for (int d = 0; d < 3; ++d) {
            maxabs = std::max(maxabs, std::max(std::fabs(bounds[d]), std::fabs(bounds[3 + d])));
            ext[d] = bounds[3 + d] - bounds[d];
            emax = std::max(emax, ext[d]);
            g.lo[d] = bounds[d];
        }
        double h = 2.0 * bounds[6] * (1.0 + 1e-6) + 1e-12 * maxabs;
        if (!(h > 0.0)) h = emax > 0.0 ? emax : 1.0;
        const uint64_t limit = std::max<uint64_t>(2ull * np, 4096);
        for (;;) {
            uint64_t total = 1;
            for (int d = 0; d < 3; ++d) {
                g.n[d] = (int)std::min(ext[d] / h, (double)(1 << 20)) + 1;
                total *= (uint64_t)g.n[d];
            }
            if (total <= limit) break;
            h *= 1.5;
        }
        g.inv_h = 1.0 / h;
        const size_t ncells = (size_t)g.n[0] * g.n[1] * g.n[2];
        AFAIL(hipMalloc(&key, (size_t)np * 4), "hipMalloc");
        AFAIL(hipMalloc(&order, (size_t)np * 4), "hipMalloc");
        AFAIL(hipMalloc(&cells, (ncells + 1) * 4), "hipMalloc");
        AFAIL(hipMalloc(&cursor, ncells * 4), "hipMalloc");
        AFAIL(hipMemsetAsync(cursor, 0, ncells * 4, s), "memset");
        AFAIL(launch_overlap_bin(x, np, g, key, cursor, s), "overlap bin");
        std::vector<uint32_t> start(ncells + 1);
        AFAIL(hipMemcpyAsync(start.data(), cursor, ncells * 4, hipMemcpyDeviceToHost, s), "D2H");
        AFAIL(hipStreamSynchronize(s), "sync");
        uint32_t run = 0;
        for (size_t k = 0; k < ncells; ++k) {
            const uint32_t n = start[k];
            start[k] = run;
            run += n;
        }
        start[ncells] = run;
        AFAIL(hipMemcpyAsync(cells, start.data(), (ncells + 1) * 4, hipMemcpyHostToDevice, s),
              "H2D");
        AFAIL(hipMemcpyAsync(cursor, cells, ncells * 4, hipMemcpyDeviceToDevice, s), "D2D");
        AFAIL(launch_overlap_scatter(key, np, cursor, order, s), "overlap scatter");
        AFAIL(hipStreamSynchronize(s), "sync");
    }
    auto search = [&](const uint64_t* o, uint32_t* p, double* dl) {
        return use_grid ? launch_overlaps_grid(x, vel, np, g, cells, order, counts, o, p, dl, s)
                        : launch_overlaps(x, vel, np, counts, o, p, dl, s);
    };
    AFAIL(search(nullptr, nullptr, nullptr), "overlaps");
    std::vector<uint32_t> hc(np);
    std::vector<uint64_t> ho(np);
    AFAIL(hipMemcpyAsync(hc.data(), counts, (size_t)np * 4, hipMemcpyDeviceToHost, s), "D2H");
    AFAIL(hipStreamSynchronize(s), "sync");
    uint64_t total = 0;
    for (uint32_t i = 0; i < np; ++i) {
        ho[i] = total;
        total += hc[i];
    }
    *count = total;
    const uint64_t take = std::min<uint64_t>(total, cap);
    if (take > 0) {
        AFAIL(hipMemcpyAsync(offs, ho.data(), (size_t)np * 8, hipMemcpyHostToDevice, s), "H2D");
        AFAIL(hipMalloc(&dpairs, total * 8), "hipMalloc");
        AFAIL(hipMalloc(&ddelta, total * 8), "hipMalloc");
        AFAIL(search(offs, dpairs, ddelta), "overlaps");
        AFAIL(hipMemcpyAsync(pairs, dpairs, take * 8, hipMemcpyDeviceToHost, s), "D2H");
        AFAIL(hipMemcpyAsync(overlaps, ddelta, take * 8, hipMemcpyDeviceToHost, s), "D2H");
    }
#undef AFAIL
    release();
    return SPH_OK;
}

int sph_porosity(sph_ctx* c, int sim, int z_slices, double* z_pos, double* porosity) {
    CHECK_SIM(c, sim);
    if (z_slices <= 0 || !z_pos || !porosity) return SPH_E_ARG;
    const Sim& S = c->sims[sim];
    const uint32_t np = S.h.np, b0 = S.P.base;
    double top = S.h.L[2];
    if (S.h.nw > 0) {
        double4 w;
        int rc = d2h(c, &w, c->A.w_nx + S.P.wall_base, 32);
        if (rc) return rc;
        top = w.w;
    }
    const double h_slice = (top - S.h.origo[2]) / (double)z_slices;
    double4 last = make_double4(0.0, 0.0, 0.0, 0.0);
    if (np > 0) {
        int rc = d2h(c, &last, c->A.x + b0 + np - 1, 32);
        if (rc) return rc;
    }
    // The reference checks the slice height against "r_max()", which
    // returns the last particle's radius (src/sphere.cpp:474-484).
    if (h_slice <= last.w * 2.0)
        return fail(c, SPH_E_INPUT, "Error! The number of z-slices is too high.");
    const double V_slice = h_slice * S.h.L[0] * S.h.L[1];
    double* d = nullptr;
    HIPCHK(c, hipMalloc(&d, (size_t)z_slices * 16), "hipMalloc");
    hipError_t e = launch_porosity(c->A.x + b0, np, z_slices, h_slice, V_slice, d, d + z_slices,
                                   c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(z_pos, d, (size_t)z_slices * 8,
                                            hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(porosity, d + z_slices, (size_t)z_slices * 8,
                                            hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(c, e, "porosity");
    return SPH_OK;
}

int sph_porosity_noalloc(sph_ctx* c, int sim, int z_slices, double* z_pos, double* porosity) {
    CHECK_SIM(c, sim);
    if (z_slices <= 0 || !z_pos || !porosity) return SPH_E_ARG;
    if (z_slices > SPH_POROSITY_NOALLOC_MAX_SLICES)
        return fail(c, SPH_E_ARG, "z_slices %d exceeds the maximum of %d for "
                    "sph_porosity_noalloc", z_slices, SPH_POROSITY_NOALLOC_MAX_SLICES);
    double* d = scratch_of(c->device);
    if (!d) return fail(c, SPH_E_DEVICE, "no staging array on device %d", c->device);
    const Sim& S = c->sims[sim];
    const uint32_t np = S.h.np, b0 = S.P.base;
    double top = S.h.L[2];
    if (S.h.nw > 0) {
        double4 w;
        int rc = d2h(c, &w, c->A.w_nx + S.P.wall_base, 32);
        if (rc) return rc;
        top = w.w;
    }
    const double h_slice = (top - S.h.origo[2]) / (double)z_slices;
    double4 last = make_double4(0.0, 0.0, 0.0, 0.0);
    if (np > 0) {
        int rc = d2h(c, &last, c->A.x + b0 + np - 1, 32);
        if (rc) return rc;
    }
    // src/sphere.cpp:474-484, as in sph_porosity.
    if (h_slice <= last.w * 2.0)
        return fail(c, SPH_E_INPUT, "Error! The number of z-slices is too high.");
    const double V_slice = h_slice * S.h.L[0] * S.h.L[1];
    // Every slice's two results are written by the kernel: nothing in the
    // array needs clearing first.
    hipError_t e = launch_porosity(c->A.x + b0, np, z_slices, h_slice, V_slice, d, d + z_slices,
                                   c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(z_pos, d, (size_t)z_slices * 8,
                                            hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(porosity, d + z_slices, (size_t)z_slices * 8,
                                            hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "porosity");
    return SPH_OK;
}

int sph_selftest_sqrt32(int device, uint64_t lo, uint64_t hi, uint64_t* mismatches) {
    if (!mismatches || lo > hi || hi > (1ull << 32)) return SPH_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return SPH_E_DEVICE;
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, sizeof *d) != hipSuccess) return SPH_E_DEVICE;
    unsigned long long n = 0;
    hipError_t e = hipMemcpy(d, &n, sizeof n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_check_sqrt32(d, lo, hi, nullptr);
    if (e == hipSuccess) e = hipMemcpy(&n, d, sizeof n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return SPH_E_DEVICE;
    *mismatches = n;
    return SPH_OK;
}

}  // extern "C"
