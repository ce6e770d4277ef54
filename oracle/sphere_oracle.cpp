// CPU oracle for the sphere DEM particle timestep -- TEST INFRASTRUCTURE.
//
// A restatement, in plain C++, of the reference's fluid-free time step.  Each
// function cites the reference code it follows.  Only tests/, the smoke()
// entry and bench.py's cpu_baseline leg load this library; the product never
// does.
//
// Arithmetic contract (the reference's, made explicit):
//  * Real is double; no FMA contraction (build with -ffp-contract=off,
//    matching nvcc --fmad=false in src/CMakeLists.txt:25).  Built with
//    -DORACLE_DUAL, Real is the dual number of dual.h (the tangent build:
//    the same values, each with a directional derivative; the orc_dual_*
//    functions seed a direction and read the derivatives).
//  * length(v) = sqrtf(dot(v,v)): the double dot is narrowed to float, the
//    float square root is correctly rounded, the result widened back
//    (src/vector_arithmetic.h:317-324).
//  * Every expression keeps the reference's evaluation order (left to
//    right; compound `a += b + c` is a = a + (b + c)).
//
// Deliberate, documented deviations from the reference:
//  * Uninitialised tangential displacement (src/contactmodels.cuh:320,405-409)
//    is defined as 0 (the projected delta_t0, which is 0 in that branch).
//  * Hertz (contactmodel 3): the reference raises a negative overlap to
//    fractional powers and yields NaN (src/contactmodels.cuh:500-559).  Here
//    |delta|^(3/2), ^(1/2), ^(1/4) are used, computed with correctly rounded
//    FP64 sqrt, and the elastic normal term is repulsive (k_n |delta|^(3/2)).
//  * bondsLinear's racy read-modify-write (src/cohesion.cuh:220-226) is
//    replaced by a serial sum in bond-index order.
//  * wall0mode for nw == 0 (read uninitialised, src/integration.cuh:57-62)
//    is taken as 0.
//  * Initial contact history: contacts = np, delta_t = 0, distmod = 0,
//    acc0 = angacc0 = 0 (the .bin does not carry them; src/file_io.cpp).

#include "sphere_oracle.h"
#include "../include/sphere_hip.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_DUAL
#include "dual.h"
#endif

namespace {

#ifdef ORACLE_DUAL
typedef Dual Real;
static inline double val(const Dual& x) { return x.v; }
#else
typedef double Real;
#endif
static inline double val(double x) { return x; }
static const int NC = SPH_NC;
static const Real PI = 3.14159265358979;  // src/constants.h:14

struct R3 { Real x, y, z; };
struct R4 { Real x, y, z, w; };

static inline R3 mk3(Real x, Real y, Real z) { R3 r = {x, y, z}; return r; }
static inline R4 mk4(Real x, Real y, Real z, Real w) { R4 r = {x, y, z, w}; return r; }
static inline R3 xyz(const R4& a) { return mk3(a.x, a.y, a.z); }

// Component-wise operators, src/vector_arithmetic.h:36-227.
static inline R3 operator+(R3 a, R3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline R3 operator-(R3 a, R3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline R3 operator+(R3 a, Real b) { return mk3(a.x + b, a.y + b, a.z + b); }   // P12
static inline R3 operator*(R3 a, Real b) { return mk3(a.x * b, a.y * b, a.z * b); }
static inline R3 operator*(Real b, R3 a) { return mk3(b * a.x, b * a.y, b * a.z); }
static inline R3 operator/(R3 a, Real b) { return mk3(a.x / b, a.y / b, a.z / b); }
static inline R3 operator-(R3 a) { return mk3(-a.x, -a.y, -a.z); }
static inline void operator+=(R3& a, R3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
static inline void operator*=(R3& a, Real b) { a.x *= b; a.y *= b; a.z *= b; }
static inline Real dot(R3 a, R3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline R3 cross(R3 a, R3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// length(): FP32 square root of the FP64 dot, src/vector_arithmetic.h:317-324.
#ifdef ORACLE_DUAL
// The FP32 value; the derivative of the exact root.
static inline Real sqrt32(Real v) {
    const double e = ::sqrt(v.v);
    return Dual((double)sqrtf((float)v.v), e > 0.0 ? v.d / (2.0 * e) : 0.0);
}
#else
static inline Real sqrt32(Real v) { return (Real)sqrtf((float)v); }
#endif
// A smooth variant for checking gradients by finite differences: lengths
// as FP64 roots (off by default; the FP32 root is the reference's).
static bool g_exact_length = false;
static bool g_wall_sums_per_wall = false;
// Test-only: record every evaluated contact's forces and displacement.
static bool g_contact_log = false;
// Record layout: active, slip, delta, n[3], f_n[3], f_t[3], f_c[3], vel_t[3],
// delta_t (stored)[3], delta_t0 (projected)[3], k_t.
static const int CLOG_W = 25;
static inline Real length(R3 v) { return g_exact_length ? sqrt(dot(v, v)) : sqrt32(dot(v, v)); }

struct Params {
    Real g[3], k_n, k_t, k_r, E, gamma_n, gamma_t, gamma_r, mu_s, mu_d, mu_r;
    Real gamma_wn, gamma_wt, mu_ws, mu_wd, rho;
    uint32_t contactmodel;
    Real kappa, db, V_b, lambda_bar;
    uint32_t nb0;
    Real sigma_b, tau_b, sigma0_A, sigma0_f;
};

}  // namespace

struct orc_sim {
    // header / grid (src/datatypes.h:45-71)
    double version;
    int32_t nd;
    uint32_t np;
    double dt, t, t_total, file_dt;
    uint32_t step_count;
    double origo[3], L[3];
    uint32_t num[3];
    int32_t periodic, adaptive;
    // particles (src/sphere.h:63-111), AoS as the reference's Float4 arrays
    std::vector<R4> x, xyzsum, vel, force, angpos, angvel, torque, acc, angacc;
    std::vector<Real> es_dot, es, ev_dot, ev, p;
    std::vector<int32_t> color;
    Params P;
    // walls (src/datatypes.h:104-110)
    uint32_t nw;
    int32_t wmode[SPH_MAXWALLS];
    std::vector<R4> w_nx, w_mvfd;
    double tau_x;
    std::vector<Real> walls_acc, force_pp, tau_pp, force_partial, tau_partial;
    // bonds
    std::vector<uint32_t> bonds;  // nb0 x 2
    std::vector<R4> bonds_delta, bonds_omega;
    // sorting (src/datatypes.h:53-61)
    std::vector<uint32_t> cellid, perm, cellStart, cellEnd;
    std::vector<R4> x_s, vel_s, angvel_s;
    // contact history
    std::vector<uint32_t> contacts;
    std::vector<R4> distmod, delta_t;
    // loop state (src/device.cu:919-993)
    int64_t iter;
    int velocity_state, change_velocity_state;
    // Test-only per-contact record of the last interact stage (models 2 and
    // 3), np x NC x CLOG_W doubles, filled when g_contact_log is set.
    std::vector<double> clog;
    // The engine's bond options (SPH_S_BOND_K, SPH_S_BOND_BREAK_EXT); 0 =
    // the reference's bondsLinear.
    double bond_k = 0.0, bond_break_ext = 0.0;

    uint32_t ncells() const { return num[0] * num[1] * num[2]; }
    uint32_t blocks() const { return (np + 255u) / 256u; }
};

namespace {

int g_threads = 0;

// ---------------------------------------------------------------------------
// .bin v2.15 reader (src/file_io.cpp:30-418, fluid == 0)

struct Reader {
    const unsigned char* p;
    size_t n, off;
    bool ok;
    template <class T> T get() {
        T v;
        if (off + sizeof(T) > n) { ok = false; memset(&v, 0, sizeof(T)); return v; }
        memcpy(&v, p + off, sizeof(T));
        off += sizeof(T);
        return v;
    }
};

struct Writer {
    std::vector<unsigned char> b;
    template <class T> void put(const T& v) {
        const unsigned char* q = reinterpret_cast<const unsigned char*>(&v);
        b.insert(b.end(), q, q + sizeof(T));
    }
#ifdef ORACLE_DUAL
    void put(const Dual& v) { put(v.v); }
#endif
};

bool parse_bin(orc_sim* s, const void* bytes, size_t nbytes, std::string& err) {
    Reader r = {static_cast<const unsigned char*>(bytes), nbytes, 0, true};
    s->version = r.get<double>();
    if (!r.ok) { err = "short read"; return false; }
    if (s->version != SPH_VERSION) {
        char m[160];
        snprintf(m, sizeof m, "input written by sphere version %g, incompatible with %g",
                 s->version, SPH_VERSION);
        err = m;
        return false;
    }
    s->nd = r.get<int32_t>();
    s->np = r.get<uint32_t>();
    if (s->nd != 3) { err = "dimensionality mismatch"; return false; }
    s->dt = r.get<double>(); s->t = r.get<double>(); s->t_total = r.get<double>();
    s->file_dt = r.get<double>(); s->step_count = r.get<uint32_t>();
    for (int d = 0; d < 3; ++d) s->origo[d] = r.get<double>();
    for (int d = 0; d < 3; ++d) s->L[d] = r.get<double>();
    for (int d = 0; d < 3; ++d) s->num[d] = r.get<uint32_t>();
    s->periodic = r.get<int32_t>(); s->adaptive = r.get<int32_t>();
    const uint32_t np = s->np;
    if ((size_t)np * 228 > nbytes) { err = "short read"; return false; }
    s->x.resize(np); s->xyzsum.resize(np); s->vel.resize(np); s->force.resize(np);
    s->angpos.resize(np); s->angvel.resize(np); s->torque.resize(np);
    for (uint32_t i = 0; i < np; ++i) {
        s->x[i].x = r.get<double>(); s->x[i].y = r.get<double>();
        s->x[i].z = r.get<double>(); s->x[i].w = r.get<double>();
    }
    for (uint32_t i = 0; i < np; ++i) {
        s->xyzsum[i].x = r.get<double>(); s->xyzsum[i].y = r.get<double>();
        s->xyzsum[i].z = r.get<double>(); s->xyzsum[i].w = 0.0;
    }
    for (uint32_t i = 0; i < np; ++i) {
        s->vel[i].x = r.get<double>(); s->vel[i].y = r.get<double>();
        s->vel[i].z = r.get<double>(); s->vel[i].w = r.get<double>();
    }
    std::vector<R4>* v3[4] = {&s->force, &s->angpos, &s->angvel, &s->torque};
    for (int a = 0; a < 4; ++a)
        for (uint32_t i = 0; i < np; ++i) {
            R4& q = (*v3[a])[i];
            q.x = r.get<double>(); q.y = r.get<double>(); q.z = r.get<double>(); q.w = 0.0;
        }
    std::vector<Real>* e[5] = {&s->es_dot, &s->es, &s->ev_dot, &s->ev, &s->p};
    for (int a = 0; a < 5; ++a) {
        e[a]->resize(np);
        for (uint32_t i = 0; i < np; ++i) (*e[a])[i] = r.get<double>();
    }
    Params& P = s->P;
    for (int d = 0; d < 3; ++d) P.g[d] = r.get<double>();
    P.k_n = r.get<double>(); P.k_t = r.get<double>(); P.k_r = r.get<double>();
    P.E = r.get<double>(); P.gamma_n = r.get<double>(); P.gamma_t = r.get<double>();
    P.gamma_r = r.get<double>(); P.mu_s = r.get<double>(); P.mu_d = r.get<double>();
    P.mu_r = r.get<double>(); P.gamma_wn = r.get<double>(); P.gamma_wt = r.get<double>();
    P.mu_ws = r.get<double>(); P.mu_wd = r.get<double>(); P.rho = r.get<double>();
    P.contactmodel = r.get<uint32_t>();
    P.kappa = r.get<double>(); P.db = r.get<double>(); P.V_b = r.get<double>();
    s->nw = r.get<uint32_t>();
    if (!r.ok) { err = "short read"; return false; }
    if (s->nw > SPH_MAXWALLS) { err = "nw exceeds MAXWALLS"; return false; }
    for (uint32_t i = 0; i < s->nw; ++i) s->wmode[i] = r.get<int32_t>();
    s->w_nx.resize(s->nw); s->w_mvfd.resize(s->nw);
    for (uint32_t i = 0; i < s->nw; ++i) {
        s->w_nx[i].x = r.get<double>(); s->w_nx[i].y = r.get<double>();
        s->w_nx[i].z = r.get<double>(); s->w_nx[i].w = r.get<double>();
    }
    for (uint32_t i = 0; i < s->nw; ++i) {
        s->w_mvfd[i].x = r.get<double>(); s->w_mvfd[i].y = r.get<double>();
        s->w_mvfd[i].z = r.get<double>(); s->w_mvfd[i].w = r.get<double>();
    }
    P.sigma0_A = r.get<double>(); P.sigma0_f = r.get<double>(); s->tau_x = r.get<double>();
    P.lambda_bar = r.get<double>(); P.nb0 = r.get<uint32_t>();
    P.sigma_b = r.get<double>(); P.tau_b = r.get<double>();
    if (!r.ok) { err = "short read"; return false; }
    const uint32_t nb = P.nb0;
    if ((size_t)nb * 72 > nbytes - r.off) { err = "short read"; return false; }
    s->bonds.resize((size_t)nb * 2);
    s->bonds_delta.assign(nb, mk4(0, 0, 0, 0));
    s->bonds_omega.assign(nb, mk4(0, 0, 0, 0));
    for (uint32_t i = 0; i < nb; ++i) {
        s->bonds[2 * i] = r.get<uint32_t>(); s->bonds[2 * i + 1] = r.get<uint32_t>();
    }
    for (uint32_t i = 0; i < nb; ++i) s->bonds_delta[i].w = r.get<double>();
    for (uint32_t i = 0; i < nb; ++i) {
        s->bonds_delta[i].x = r.get<double>(); s->bonds_delta[i].y = r.get<double>();
        s->bonds_delta[i].z = r.get<double>();
    }
    for (uint32_t i = 0; i < nb; ++i) s->bonds_omega[i].w = r.get<double>();
    for (uint32_t i = 0; i < nb; ++i) {
        s->bonds_omega[i].x = r.get<double>(); s->bonds_omega[i].y = r.get<double>();
        s->bonds_omega[i].z = r.get<double>();
    }
    s->color.resize(np);
    for (uint32_t i = 0; i < np; ++i) s->color[i] = r.get<int32_t>();
    if (!r.ok) { err = "short read"; return false; }
    return true;
}

void write_bin(const orc_sim* s, Writer& w) {
    w.put(s->version); w.put(s->nd); w.put(s->np);
    w.put(s->dt); w.put(s->t); w.put(s->t_total); w.put(s->file_dt); w.put(s->step_count);
    for (int d = 0; d < 3; ++d) w.put(s->origo[d]);
    for (int d = 0; d < 3; ++d) w.put(s->L[d]);
    for (int d = 0; d < 3; ++d) w.put(s->num[d]);
    w.put(s->periodic); w.put(s->adaptive);
    const uint32_t np = s->np;
    for (uint32_t i = 0; i < np; ++i) { w.put(s->x[i].x); w.put(s->x[i].y); w.put(s->x[i].z); w.put(s->x[i].w); }
    for (uint32_t i = 0; i < np; ++i) { w.put(s->xyzsum[i].x); w.put(s->xyzsum[i].y); w.put(s->xyzsum[i].z); }
    for (uint32_t i = 0; i < np; ++i) { w.put(s->vel[i].x); w.put(s->vel[i].y); w.put(s->vel[i].z); w.put(s->vel[i].w); }
    const std::vector<R4>* v3[4] = {&s->force, &s->angpos, &s->angvel, &s->torque};
    for (int a = 0; a < 4; ++a)
        for (uint32_t i = 0; i < np; ++i) { w.put((*v3[a])[i].x); w.put((*v3[a])[i].y); w.put((*v3[a])[i].z); }
    const std::vector<Real>* e[5] = {&s->es_dot, &s->es, &s->ev_dot, &s->ev, &s->p};
    for (int a = 0; a < 5; ++a)
        for (uint32_t i = 0; i < np; ++i) w.put((*e[a])[i]);
    const Params& P = s->P;
    for (int d = 0; d < 3; ++d) w.put(P.g[d]);
    w.put(P.k_n); w.put(P.k_t); w.put(P.k_r); w.put(P.E); w.put(P.gamma_n); w.put(P.gamma_t);
    w.put(P.gamma_r); w.put(P.mu_s); w.put(P.mu_d); w.put(P.mu_r); w.put(P.gamma_wn);
    w.put(P.gamma_wt); w.put(P.mu_ws); w.put(P.mu_wd); w.put(P.rho); w.put(P.contactmodel);
    w.put(P.kappa); w.put(P.db); w.put(P.V_b);
    w.put(s->nw);
    for (uint32_t i = 0; i < s->nw; ++i) w.put(s->wmode[i]);
    for (uint32_t i = 0; i < s->nw; ++i) { w.put(s->w_nx[i].x); w.put(s->w_nx[i].y); w.put(s->w_nx[i].z); w.put(s->w_nx[i].w); }
    for (uint32_t i = 0; i < s->nw; ++i) { w.put(s->w_mvfd[i].x); w.put(s->w_mvfd[i].y); w.put(s->w_mvfd[i].z); w.put(s->w_mvfd[i].w); }
    w.put(P.sigma0_A); w.put(P.sigma0_f); w.put(s->tau_x);
    w.put(P.lambda_bar); w.put(P.nb0); w.put(P.sigma_b); w.put(P.tau_b);
    for (uint32_t i = 0; i < P.nb0; ++i) { w.put(s->bonds[2 * i]); w.put(s->bonds[2 * i + 1]); }
    for (uint32_t i = 0; i < P.nb0; ++i) w.put(s->bonds_delta[i].w);
    for (uint32_t i = 0; i < P.nb0; ++i) { w.put(s->bonds_delta[i].x); w.put(s->bonds_delta[i].y); w.put(s->bonds_delta[i].z); }
    for (uint32_t i = 0; i < P.nb0; ++i) w.put(s->bonds_omega[i].w);
    for (uint32_t i = 0; i < P.nb0; ++i) { w.put(s->bonds_omega[i].x); w.put(s->bonds_omega[i].y); w.put(s->bonds_omega[i].z); }
    for (uint32_t i = 0; i < np; ++i) w.put(s->color[i]);
}

void init_runtime(orc_sim* s) {
    const uint32_t np = s->np;
    s->acc.assign(np, mk4(0, 0, 0, 0));
    s->angacc.assign(np, mk4(0, 0, 0, 0));
    s->cellid.assign(np, 0);
    s->perm.assign(np, 0);
    s->cellStart.assign(s->ncells(), 0xffffffffu);
    s->cellEnd.assign(s->ncells(), 0u);
    s->x_s.assign(np, mk4(0, 0, 0, 0));
    s->vel_s.assign(np, mk4(0, 0, 0, 0));
    s->angvel_s.assign(np, mk4(0, 0, 0, 0));
    // src/device.cu:458-462: contact lists start at np; delta_t and distmod are
    // left uninitialised there and are defined as zero here.
    s->contacts.assign((size_t)np * NC, np);
    s->distmod.assign((size_t)np * NC, mk4(0, 0, 0, 0));
    s->delta_t.assign((size_t)np * NC, mk4(0, 0, 0, 0));
    s->walls_acc.assign(s->nw, 0.0);
    s->force_pp.assign((size_t)s->nw * np, 0.0);
    s->tau_pp.assign((size_t)s->nw * np, 0.0);
    s->force_partial.assign(s->blocks(), 0.0);
    s->tau_partial.assign(s->blocks(), 0.0);
    s->iter = 0;
    s->velocity_state = 1;
    s->change_velocity_state = 0;
}

// ---------------------------------------------------------------------------
// Hashing and sorting

// calcCellID, src/sorting.cuh:7-20.
inline uint32_t calc_cell_id(const orc_sim* s, Real x, Real y, Real z) {
    uint32_t ix = (uint32_t)floor((x - s->origo[0]) / (s->L[0] / s->num[0]));
    uint32_t iy = (uint32_t)floor((y - s->origo[1]) / (s->L[1] / s->num[1]));
    uint32_t iz = (uint32_t)floor((z - s->origo[2]) / (s->L[2] / s->num[2]));
    return (iz * s->num[1]) * s->num[0] + iy * s->num[0] + ix;
}

// calcParticleCellID, src/sorting.cuh:25-51.
void stage_hash(orc_sim* s) {
    const int64_t np = s->np;
#pragma omp parallel for schedule(static) if (np > 4096)
    for (int64_t i = 0; i < np; ++i) {
        s->cellid[i] = calc_cell_id(s, s->x[i].x, s->x[i].y, s->x[i].z);
        s->perm[i] = (uint32_t)i;
    }
}

// thrust::sort_by_key(cellID, index), src/device.cu:1075-1078.  Its radix
// path is stable, so the result is the lexicographic order of
// (cellID, original index) -- a stable counting sort gives exactly that.
void stage_sort(orc_sim* s) {
    const uint32_t np = s->np;
    std::vector<uint32_t> key(s->cellid), idx(s->perm);
    std::vector<uint32_t> order(np);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
    for (uint32_t i = 0; i < np; ++i) {
        s->cellid[i] = key[order[i]];
        s->perm[i] = idx[order[i]];
    }
}

// cudaMemset(cellStart, 0xff) + reorderArrays, src/device.cu:1089-1090 and
// src/sorting.cuh:56-132.  cellEnd is never cleared (P16).
void stage_reorder(orc_sim* s) {
    std::fill(s->cellStart.begin(), s->cellStart.end(), 0xffffffffu);
    const uint32_t np = s->np;
    for (uint32_t i = 0; i < np; ++i) {
        // A key >= ncells (a particle at exactly x == L under the periodic
        // wrap, P17) has no cell entry: the reference writes out of bounds
        // there; those writes are skipped.
        const uint32_t c = s->cellid[i];
        const uint32_t nc = s->ncells();
        if (i == 0 || c != s->cellid[i - 1]) {
            if (c < nc) s->cellStart[c] = i;
            if (i > 0 && s->cellid[i - 1] < nc) s->cellEnd[s->cellid[i - 1]] = i;
        }
        if (i == np - 1 && c < nc) s->cellEnd[c] = i + 1;
    }
#pragma omp parallel for schedule(static) if (np > 4096)
    for (int64_t i = 0; i < (int64_t)np; ++i) {
        const uint32_t o = s->perm[i];
        s->x_s[i] = s->x[o];
        s->vel_s[i] = s->vel[o];
        s->angvel_s[i] = s->angvel[o];
    }
}

// ---------------------------------------------------------------------------
// Contact search

// findDistMod, src/contactsearch.cuh:12-83.
inline int find_dist_mod(const orc_sim* s, int tc[3], R3* dm) {
    const int nx = (int)s->num[0], ny = (int)s->num[1], nz = (int)s->num[2];
    if (s->periodic == 1) {
        if (tc[0] < 0) { tc[0] += nx; *dm += mk3(s->L[0], 0.0, 0.0); }
        if (tc[0] >= nx) { tc[0] -= nx; *dm = *dm - mk3(s->L[0], 0.0, 0.0); }
        if (tc[1] < 0) { tc[1] += ny; *dm += mk3(0.0, s->L[1], 0.0); }
        if (tc[1] >= ny) { tc[1] -= ny; *dm = *dm - mk3(0.0, s->L[1], 0.0); }
    }
    if (s->periodic == 2) {
        if (tc[0] < 0) { tc[0] += nx; *dm += mk3(s->L[0], 0.0, 0.0); }
        if (tc[0] >= nx) { tc[0] -= nx; *dm = *dm - mk3(s->L[0], 0.0, 0.0); }
        if (tc[1] < 0 || tc[1] >= ny) return -1;
    }
    if (s->periodic > 2 || s->periodic < 1) {
        if (tc[0] < 0 || tc[0] >= nx) return -1;
        if (tc[1] < 0 || tc[1] >= ny) return -1;
    }
    if (tc[2] < 0 || tc[2] >= nz) return -1;
    return 0;
}

inline void grid_pos(const orc_sim* s, const R4& xa, int g[3]) {
    // src/contactsearch.cuh:352-354 (double floor converted to int)
    g[0] = (int)floor((xa.x - s->origo[0]) / (s->L[0] / s->num[0]));
    g[1] = (int)floor((xa.y - s->origo[1]) / (s->L[1] / s->num[1]));
    g[2] = (int)floor((xa.z - s->origo[2]) / (s->L[2] / s->num[2]));
}

// findContactsInCell, src/contactsearch.cuh:194-322, including the
// contact-slot policy (P5).  Returns after the overlap that makes *nc
// exceed NC, ending the scan of this cell only.
void find_contacts_in_cell(orc_sim* s, int tc[3], uint32_t ia, const R4& xa, int* nc) {
    R3 dm = mk3(0.0, 0.0, 0.0);
    if (find_dist_mod(s, tc, &dm) == -1) return;
    const uint32_t cell = (uint32_t)tc[0] + (uint32_t)tc[1] * s->num[0] +
                          (s->num[0] * s->num[1]) * (uint32_t)tc[2];
    const uint32_t start = s->cellStart[cell];
    if (start == 0xffffffffu) return;
    const uint32_t end = s->cellEnd[cell];
    const uint32_t ia_orig = s->perm[ia];
    const Real ra = xa.w;
    for (uint32_t ib = start; ib < end; ++ib) {
        if (ib == ia) continue;
        const R4 xb = s->x_s[ib];
        const Real rb = xb.w;
        const uint32_t ib_orig = s->perm[ib];
        R3 xab = mk3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z);
        xab += dm;
        const Real len = length(xab);
        const Real delta = len - (ra + rb);
        if (delta < 0.0) {
            int cpos = *nc;
            for (int i = 0; i < NC; ++i) {
                const uint32_t cidx = s->contacts[(size_t)ia_orig * NC + i];
                if (cidx == s->np) {
                    cpos = i;
                } else if (cidx == ib_orig) {
                    cpos = i;
                    break;
                }
            }
            // cpos reaches NC only through *nc == NC with a full row that
            // holds neither an empty slot nor B.  The reference then writes
            // slot 0 of the next particle's row (src/contactsearch.cuh:286),
            // a cross-row race; that one write is suppressed here.
            if (cpos < NC) {
                const size_t mp = (size_t)ia_orig * NC + (size_t)cpos;
                s->contacts[mp] = ib_orig;
                s->distmod[mp] = mk4(dm.x, dm.y, dm.z, rb);
            }
            ++*nc;
            if (*nc > NC) return;
        }
    }
}

// topology, src/contactsearch.cuh:329-371.
void stage_topology(orc_sim* s) {
    const int64_t np = s->np;
#pragma omp parallel for schedule(dynamic, 256) if (np > 1024)
    for (int64_t ia = 0; ia < np; ++ia) {
        const R4 xa = s->x_s[ia];
        int nc = 0;
        int g[3];
        grid_pos(s, xa, g);
        for (int dz = -1; dz < 2; ++dz)
            for (int dy = -1; dy < 2; ++dy)
                for (int dx = -1; dx < 2; ++dx) {
                    int tc[3] = {g[0] + dx, g[1] + dy, g[2] + dz};
                    find_contacts_in_cell(s, tc, (uint32_t)ia, xa, &nc);
                }
    }
}

// ---------------------------------------------------------------------------
// Contact models (src/contactmodels.cuh)

struct Acc {
    R3 F, T;
    Real es_dot, ev_dot, p;
};

// contactLinear_wall, src/contactmodels.cuh:10-117.
Real contact_linear_wall(const orc_sim* s, Acc& A, uint32_t ia, Real ra, R3 n, Real delta,
                         Real wvel) {
    const Params& P = s->P;
    const R4 vt = s->vel_s[ia];
    const R4 at = s->angvel_s[ia];
    const R3 vel_linear = xyz(vt);
    const R3 angvel = xyz(at);
    const R3 vel = vel_linear + (ra + delta / 2.0) * cross(n, angvel) + wvel;
    const Real vel_n = dot(vel_linear, n);
    const R3 vel_t = vel - n * (dot(n, vel));
    const Real vel_t_length = length(vel_t);
    Real k_n = P.k_n;
    if (P.E > .001) k_n = 3.141592654 / 2.0 * P.E * ra;
    const R3 f_n = fmax(0.0, -k_n * delta - P.gamma_wn * vel_n) * n;
    const Real f_n_length = length(f_n);
    A.ev_dot += P.gamma_wn * vel_n * vel_n;
    R3 f_t = mk3(0.0, 0.0, 0.0);
    if (vel_t_length > 0.0 && P.gamma_wt > 0.0) {
        const R3 f_t_visc = P.gamma_wt * vel_t;
        const Real f_t_visc_length = length(f_t_visc);
        // vel_t_length > 0 always holds here: the dynamic limit applies
        // (src/contactmodels.cuh:86-90).
        const Real f_t_limit = P.mu_wd * f_n_length;
        if (f_t_visc_length < f_t_limit) {
            f_t = -1.0 * f_t_visc;
        } else {
            f_t = -f_t_limit * vel_t / vel_t_length;
            A.es_dot += length(length(f_t) * vel_t * s->dt) / s->dt;
        }
    }
    A.F += f_n + f_t;
    A.T += cross(-(ra + delta * 0.5) * n, f_t);
    A.p += f_n_length / (4.0 * PI * ra * ra);
    return dot(f_n, n);
}

// contactLinearViscous (model 1), src/contactmodels.cuh:122-240.
void contact_linear_viscous(const orc_sim* s, Acc& A, uint32_t ia, uint32_t ib, Real ra,
                            Real rb, R3 x_ab, Real x_ab_length, Real delta, Real kappa) {
    const Params& P = s->P;
    const R4 vel_a = s->vel_s[ia], vel_b = s->vel_s[ib];
    const R3 angvel_a = xyz(s->angvel_s[ia]), angvel_b = xyz(s->angvel_s[ib]);
    R3 f_n, f_t, f_c;
    const R3 n_ab = x_ab / x_ab_length;
    const R3 vel_ab_linear = mk3(vel_a.x - vel_b.x, vel_a.y - vel_b.y, vel_a.z - vel_b.z);
    const R3 vel_ab = vel_ab_linear + (ra + delta / 2.0) * cross(n_ab, angvel_a) +
                      (rb + delta / 2.0) * cross(n_ab, angvel_b);
    const Real vel_n_ab = dot(vel_ab_linear, n_ab);
    const R3 vel_t_ab = vel_ab - (n_ab * dot(vel_ab, n_ab));
    const Real vel_t_ab_length = length(vel_t_ab);
    Real k_n = P.k_n;
    if (P.E > .001) k_n = 3.141592654 / 2.0 * P.E * (ra + rb) / 2.;
    f_n = fmax(0.0, -k_n * delta - P.gamma_n * vel_n_ab) * n_ab;
    if (dot(f_n, n_ab) < 0.0) f_n = mk3(0.0, 0.0, 0.0);
    const Real f_n_length = length(f_n);
    f_c = -kappa * sqrt32(ra * rb) * n_ab;
    f_t = mk3(0.0, 0.0, 0.0);
    if (vel_t_ab_length > 0.0) {
        const Real f_t_visc = P.gamma_t * vel_t_ab_length;
        Real f_t_limit;
        if (vel_t_ab_length > (Real)0.001f)
            f_t_limit = P.mu_d * length(f_n - f_c);
        else
            f_t_limit = P.mu_s * length(f_n - f_c);
        if (f_t_visc < f_t_limit)
            f_t = -f_t_visc * vel_t_ab / vel_t_ab_length;
        else
            f_t = -f_t_limit * vel_t_ab / vel_t_ab_length;
    }
    A.F += f_n + f_t + f_c;
    A.T += -(ra + delta / 2.0) * cross(n_ab, f_t);
    A.p += f_n_length / (4.0 * PI * ra * ra);
}

// capillaryCohesion_exp, src/cohesion.cuh:237-277 (FP32 sqrtf/expf as there).
void capillary_cohesion_exp(const orc_sim* s, Acc& A, Real ra, Real rb, Real delta, R3 x_ab,
                            Real x_ab_length, Real kappa) {
    const R3 n_ab = x_ab / x_ab_length;
    Real r;
    if ((ra / rb) > (rb / ra)) r = ra / rb; else r = rb / ra;
    const Real h = -sqrt32(r);
    const Real R_har = (2.0 * ra * rb) / (ra + rb);
    const Real R_geo = sqrt32(ra * rb);
    const Real lambda = 0.9 * h * sqrt32(s->P.V_b / R_har);
    const R3 f_c = -kappa * R_geo * (Real)expf((float)(-delta / lambda)) * n_ab;
    A.F += f_c;
}

static inline void log_put3(double* r, int at, const R3& v) {
    r[at] = val(v.x); r[at + 1] = val(v.y); r[at + 2] = val(v.z);
}

// Test-only record of one evaluated contact (see CLOG_W).
static inline void log_contact(orc_sim* s, size_t mempos, bool slip, Real delta, R3 n, R3 f_n,
                               R3 f_t, R3 f_c, R3 vel_t, R3 delta_t, R3 delta_t0, Real k_t) {
    if (!g_contact_log) return;
    double* r = &s->clog[mempos * CLOG_W];
    r[0] = 1.0;
    r[1] = slip ? 1.0 : 0.0;
    r[2] = val(delta);
    log_put3(r, 3, n);
    log_put3(r, 6, f_n);
    log_put3(r, 9, f_t);
    log_put3(r, 12, f_c);
    log_put3(r, 15, vel_t);
    log_put3(r, 18, delta_t);
    log_put3(r, 21, delta_t0);
    r[24] = val(k_t);
}

// contactLinear (model 2), src/contactmodels.cuh:244-411.
void contact_linear(orc_sim* s, Acc& A, uint32_t ib_orig, const R4& vel_a, R3 angvel_a,
                    Real ra, Real rb, R3 x, Real x_length, Real delta, size_t mempos) {
    const Params& P = s->P;
    const R4 vel_b = s->vel[ib_orig];
    const R4 av_b = s->angvel[ib_orig];
    const R4 dt0 = s->delta_t[mempos];
    const R3 delta_t0_uncor = mk3(dt0.x, dt0.y, dt0.z);
    const R3 angvel_b = xyz(av_b);
    R3 f_n, f_t, f_c;
    const R3 n = x / x_length;
    const R3 vel_linear = mk3(vel_a.x - vel_b.x, vel_a.y - vel_b.y, vel_a.z - vel_b.z);
    const R3 vel = vel_linear + (ra + delta / 2.0) * cross(n, angvel_a) +
                   (rb + delta / 2.0) * cross(n, angvel_b);
    const Real vel_n = -dot(vel_linear, n);
    const R3 vel_t = vel - n * dot(n, vel);
    const Real vel_t_length = length(vel_t);
    const R3 delta_t0 = delta_t0_uncor - (n * dot(n, delta_t0_uncor));
    const Real delta_t0_length = length(delta_t0);
    R3 delta_t = delta_t0;  // P4: the stored value when neither branch assigns
    Real k_n = P.k_n, k_t = P.k_t;
    if (P.E > .001) {
        k_n = 3.141592654 / 2.0 * P.E * (ra + rb) / 2.;
        k_t = k_n;
    }
    f_n = fmax(0.0, -k_n * delta + P.gamma_n * vel_n) * n;
    const Real f_n_length = length(f_n);
    A.ev_dot += 0.5 * P.gamma_n * vel_n * vel_n;
    f_c = -P.kappa * sqrt32(ra * rb) * n;
    f_t = mk3(0.0, 0.0, 0.0);
    bool slip = false;
    if (delta_t0_length > 0.0 || vel_t_length > 0.0) {
        delta_t = delta_t0 + vel_t * s->dt;
        const R3 f_t_elast = -k_t * delta_t;
        const R3 f_t_visc = -P.gamma_t * vel_t;
        f_t = f_t_elast + f_t_visc;
        const Real f_t_length = length(f_t);
        const Real f_t_limit = P.mu_s * length(f_n - f_c);
        A.ev_dot += 0.5 * P.gamma_t * vel_t_length * vel_t_length;
        if (f_t_length > f_t_limit) {
            slip = true;
            const R3 t = f_t / length(f_t);
            f_t = f_t_limit * t;
            delta_t = -1.0 / k_t * (P.mu_d * length(f_n - f_c) * t + P.gamma_t * vel_t);
            A.es_dot += 0.5 * length(length(f_t) * vel_t * s->dt) / s->dt;
        }
    }
    A.F += f_n + f_t + f_c;
    A.T += cross(-(ra + delta * 0.5) * n, f_t);
    A.p += f_n_length / (4.0 * PI * ra * ra);
    s->delta_t[mempos] = mk4(delta_t.x, delta_t.y, delta_t.z, 0.0);
    log_contact(s, mempos, slip, delta, n, f_n, f_t, f_c, vel_t, delta_t, delta_t0, k_t);
}

// contactHertz (model 3), src/contactmodels.cuh:417-600, with the documented
// |delta| deviation (P3): powers of |delta| through correctly rounded sqrt.
void contact_hertz(orc_sim* s, Acc& A, uint32_t ib_orig, const R4& vel_a, R3 angvel_a,
                   Real ra, Real rb, R3 x_ab, Real x_ab_length, Real delta, size_t mempos) {
    const Params& P = s->P;
    const R4 vel_b = s->vel[ib_orig];
    const R3 angvel_b = xyz(s->angvel[ib_orig]);
    const R4 dt0 = s->delta_t[mempos];
    const R3 delta_t0_uncor = mk3(dt0.x, dt0.y, dt0.z);
    R3 f_n, f_t, f_c, T_res;
    const R3 n_ab = x_ab / x_ab_length;
    const R3 vel_ab_linear = mk3(vel_a.x - vel_b.x, vel_a.y - vel_b.y, vel_a.z - vel_b.z);
    const R3 vel_ab = vel_ab_linear + (ra + delta / 2.0) * cross(n_ab, angvel_a) +
                      (rb + delta / 2.0) * cross(n_ab, angvel_b);
    const R3 angvel_ab = angvel_a - angvel_b;
    const Real angvel_ab_length = length(angvel_ab);
    const Real vel_n_ab = dot(vel_ab_linear, n_ab);
    const R3 vel_t_ab = vel_ab - (n_ab * dot(vel_ab, n_ab));
    const Real vel_t_ab_length = length(vel_t_ab);
    const R3 delta_t0 = delta_t0_uncor - (n_ab * dot(delta_t0_uncor, n_ab));
    const Real delta_t0_length = length(delta_t0);
    R3 delta_t = delta_t0;
    Real k_n = P.k_n, k_t = P.k_t;
    if (P.E > .001) {
        k_n = 3.141592654 / 2.0 * P.E * (ra + rb) / 2.;
        k_t = k_n;
    }
    const Real ad = fabs(delta);
    const Real p12 = sqrt(ad);
    const Real p14 = sqrt(p12);
    const Real p32 = ad * p12;
    f_n = (k_n * p32 - P.gamma_n * p14 * vel_n_ab) * n_ab;
    A.ev_dot += P.gamma_n * vel_n_ab * vel_n_ab;
    if (dot(f_n, n_ab) < 0.0) f_n = mk3(0.0, 0.0, 0.0);
    const Real f_n_length = length(f_n);
    f_c = -P.kappa * sqrt32(ra * rb) * n_ab;
    f_t = mk3(0.0, 0.0, 0.0);
    T_res = mk3(0.0, 0.0, 0.0);
    bool slip = false;
    if (delta_t0_length > 0.0 || vel_t_ab_length > 0.0) {
        const R3 f_t_elast = -k_t * p12 * delta_t0;
        const R3 f_t_visc = -P.gamma_t * p14 * vel_t_ab;
        Real f_t_limit;
        if (vel_t_ab_length > (Real)0.001f)
            f_t_limit = P.mu_d * length(f_n - f_c);
        else
            f_t_limit = P.mu_s * length(f_n - f_c);
        f_t = f_t_elast + f_t_visc;
        const Real f_t_length = length(f_t);
        if (f_t_length > f_t_limit) {
            slip = true;
            f_t *= f_t_limit / f_t_length;
            delta_t = (f_t + P.gamma_t * p14 * vel_t_ab) / (-k_t * p12);
            A.es_dot += length(delta_t0 - delta_t) * k_t / s->dt;
        } else {
            delta_t = delta_t0 + vel_t_ab * s->dt;
        }
    }
    if (angvel_ab_length > 0.0) {
        T_res = -1.0 * fmin(P.gamma_r * ra * angvel_ab_length, P.mu_r * ra * f_n_length) *
                angvel_ab / angvel_ab_length;
    }
    A.F += f_n + f_t + f_c;
    A.T += -(ra + delta / 2.0) * cross(n_ab, f_t) + T_res;
    A.p += f_n_length / (4.0 * PI * ra * ra);
    s->delta_t[mempos] = mk4(delta_t.x, delta_t.y, delta_t.z, 0.0);
    log_contact(s, mempos, slip, delta, n_ab, f_n, f_t, f_c, vel_t_ab, delta_t, delta_t0, k_t);
}

// findAndProcessContactsInCell (model 1), src/contactsearch.cuh:91-186.
void find_and_process_in_cell(orc_sim* s, Acc& A, int tc[3], uint32_t ia, const R4& xa) {
    R3 dm = mk3(0.0, 0.0, 0.0);
    if (find_dist_mod(s, tc, &dm) == -1) return;
    const uint32_t cell = (uint32_t)tc[0] + (uint32_t)tc[1] * s->num[0] +
                          (s->num[0] * s->num[1]) * (uint32_t)tc[2];
    const uint32_t start = s->cellStart[cell];
    if (start == 0xffffffffu) return;
    const uint32_t end = s->cellEnd[cell];
    const Real ra = xa.w;
    for (uint32_t ib = start; ib < end; ++ib) {
        if (ib == ia) continue;
        const R4 xb = s->x_s[ib];
        const Real rb = xb.w;
        const Real kappa = s->P.kappa;
        R3 xab = mk3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z);
        xab += dm;
        const Real len = length(xab);
        const Real delta = len - (ra + rb);
        if (delta < 0.0)
            contact_linear_viscous(s, A, ia, ib, ra, rb, xab, len, delta, kappa);
        else if (delta < s->P.db)
            capillary_cohesion_exp(s, A, ra, rb, delta, xab, len, kappa);
    }
}

// interact, src/contactsearch.cuh:385-726.
void stage_interact(orc_sim* s) {
    const int64_t np = s->np;
    const uint32_t nw = s->nw;
    const R3 origo = mk3(s->origo[0], s->origo[1], s->origo[2]);
    const R3 L = mk3(s->L[0], s->L[1], s->L[2]);
    R4 wn[5] = {mk4(0.0, 0.0, -1.0, L.z), mk4(-1.0, 0.0, 0.0, L.x), mk4(1.0, 0.0, 0.0, 0.0),
                mk4(0.0, -1.0, 0.0, L.y), mk4(0.0, 1.0, 0.0, 0.0)};
    R4 wm[5] = {mk4(0, 0, 0, 0), mk4(0, 0, 0, 0), mk4(0, 0, 0, 0), mk4(0, 0, 0, 0),
                mk4(0, 0, 0, 0)};
    for (uint32_t w = 0; w < nw && w < 5; ++w) { wn[w] = s->w_nx[w]; wm[w] = s->w_mvfd[w]; }
    const uint32_t cm = s->P.contactmodel;
    if (g_contact_log) s->clog.assign((size_t)np * NC * CLOG_W, 0.0);
#pragma omp parallel for schedule(dynamic, 256) if (np > 1024)
    for (int64_t ia64 = 0; ia64 < np; ++ia64) {
        const uint32_t ia = (uint32_t)ia64;
        const uint32_t ia_orig = s->perm[ia];
        const R4 xa = s->x_s[ia];
        const Real ra = xa.w;
        Acc A;
        A.es_dot = 0.0; A.ev_dot = 0.0; A.p = 0.0;
        A.F = mk3(0.0, 0.0, 0.0); A.T = mk3(0.0, 0.0, 0.0);
        if (cm == 2 || cm == 3) {
            const R4 vel_a = s->vel_s[ia];
            const R3 angvel_a = xyz(s->angvel_s[ia]);
            for (int i = 0; i < NC; ++i) {
                const size_t mp = (size_t)ia_orig * NC + i;
                const uint32_t ib_orig = s->contacts[mp];
                if (ib_orig != s->np) {
                    const R4 dm = s->distmod[mp];
                    const R4 xb = s->x[ib_orig];
                    const Real rb = xb.w;
                    const R3 xab = mk3(xa.x - xb.x + dm.x, xa.y - xb.y + dm.y, xa.z - xb.z + dm.z);
                    const Real len = length(xab);
                    const Real dn = len - (ra + rb);
                    if (dn < 0.0) {
                        if (cm == 2)
                            contact_linear(s, A, ib_orig, vel_a, angvel_a, ra, rb, xab, len, dn, mp);
                        else
                            contact_hertz(s, A, ib_orig, vel_a, angvel_a, ra, rb, xab, len, dn, mp);
                    } else {
                        s->contacts[mp] = s->np;
                        s->delta_t[mp] = mk4(0.0, 0.0, 0.0, 0.0);
                    }
                } else {
                    s->delta_t[mp] = mk4(0.0, 0.0, 0.0, 0.0);
                }
            }
        } else if (cm == 1) {
            int g[3];
            grid_pos(s, xa, g);
            for (int dz = -1; dz < 2; ++dz)
                for (int dy = -1; dy < 2; ++dy)
                    for (int dx = -1; dx < 2; ++dx) {
                        int tc[3] = {g[0] + dx, g[1] + dy, g[2] + dz};
                        find_and_process_in_cell(s, A, tc, ia, xa);
                    }
        }
        // Walls, src/contactsearch.cuh:602-701.
        Real wf[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        Real dw = wn[0].w - (xa.z + ra);
        if (dw < 0.0) wf[0] = contact_linear_wall(s, A, ia, ra, xyz(wn[0]), dw, wm[0].y);
        dw = xa.z - ra - origo.z;
        if (dw < 0.0) (void)contact_linear_wall(s, A, ia, ra, mk3(0.0, 0.0, 1.0), dw, 0.0);
        if (s->periodic == 0) {
            dw = wn[1].w - (xa.x + ra);
            if (dw < 0.0) wf[1] = contact_linear_wall(s, A, ia, ra, xyz(wn[1]), dw, wm[1].y);
            dw = xa.x - ra - wn[2].w;
            if (dw < 0.0) wf[2] = contact_linear_wall(s, A, ia, ra, xyz(wn[2]), dw, wm[2].y);
            dw = wn[3].w - (xa.y + ra);
            if (dw < 0.0) wf[3] = contact_linear_wall(s, A, ia, ra, xyz(wn[3]), dw, wm[3].y);
            dw = xa.y - ra - wn[4].w;
            if (dw < 0.0) wf[4] = contact_linear_wall(s, A, ia, ra, xyz(wn[4]), dw, wm[4].y);
        } else if (s->periodic == 2) {
            dw = wn[3].w - (xa.y + ra);
            if (dw < 0.0) wf[3] = contact_linear_wall(s, A, ia, ra, xyz(wn[3]), dw, wm[3].y);
            dw = xa.y - ra - wn[4].w;
            if (dw < 0.0) wf[4] = contact_linear_wall(s, A, ia, ra, xyz(wn[4]), dw, wm[4].y);
        }
        const uint32_t o = ia_orig;
        s->force[o] = mk4(A.F.x, A.F.y, A.F.z, 0.0);
        s->torque[o] = mk4(A.T.x, A.T.y, A.T.z, 0.0);
        s->es_dot[o] = A.es_dot;
        s->ev_dot[o] = A.ev_dot;
        s->es[o] += A.es_dot * s->dt;
        s->ev[o] += A.ev_dot * s->dt;
        s->p[o] = A.p;
        for (uint32_t w = 0; w < nw && w < 5; ++w) s->force_pp[(size_t)w * np + o] = wf[w];
    }
}

// ---------------------------------------------------------------------------
// Bonds: bondsLinear, src/cohesion.cuh:8-233, summed serially in bond order.
void stage_bonds(orc_sim* s) {
    const Params& P = s->P;
    const Real dt = s->dt;
    for (uint32_t b = 0; b < P.nb0; ++b) {
        const uint32_t bx = s->bonds[2 * b], by = s->bonds[2 * b + 1];
        if (!(bx < s->np)) continue;
        const R4 d0 = s->bonds_delta[b], o0 = s->bonds_omega[b];
        R3 delta0_t = mk3(d0.x, d0.y, d0.z);
        const Real delta0_n = d0.w;
        R3 omega0_t = mk3(o0.x, o0.y, o0.z);
        const Real omega0_n = o0.w;
        const R4 x_i = s->x[bx], x_j = s->x[by];
        const R4 vel_i = s->vel[bx], vel_j = s->vel[by];
        const R3 angvel_i = xyz(s->angvel[bx]), angvel_j = xyz(s->angvel[by]);
        const Real R_bar = P.lambda_bar * fmin(x_i.w, x_j.w);
        const Real A = PI * R_bar * R_bar;
        const Real I = 0.25 * PI * R_bar * R_bar * R_bar * R_bar;
        const Real J = I * 2.0;
        const R3 x = mk3(x_i.x - x_j.x, x_i.y - x_j.y, x_i.z - x_j.z);
        const Real x_length = length(x);
        const Real overlap = fmin(0.0, x_length - (x_i.w + x_j.w));
        const R3 n = x / x_length;
        delta0_t = delta0_t - (n * dot(n, delta0_t));
        const R3 ddelta = (mk3(vel_i.x - vel_j.x, vel_i.y - vel_j.y, vel_i.z - vel_j.z) +
                           (x_i.w + overlap / 2.0) * cross(n, angvel_i) +
                           (x_j.w + overlap / 2.0) * cross(n, angvel_j)) * dt;
        const Real ddelta_n = -dot(ddelta, n);
        const Real delta_n = delta0_n + ddelta_n;
        const R3 ddelta_t = ddelta - n * dot(n, ddelta);
        const R3 delta_t = delta0_t + ddelta_t;
        // Stiffness per unit area: k_n, k_t, or the engine's bond option
        // (a stiffness in N/m, bond_k / A per unit area).
        const Real kn = s->bond_k > 0.0 ? s->bond_k / A : P.k_n;
        const Real kt = s->bond_k > 0.0 ? s->bond_k / A : P.k_t;
        const R3 f_n = (kn * A * delta_n + P.gamma_n * ddelta_n / dt) * n;
        const R3 f_t = -kt * A * delta_t - P.gamma_t * ddelta_t / dt;
        const R3 f = f_n + f_t;
        const R3 mn = -n;
        omega0_t = omega0_t - mn * dot(omega0_t, mn);
        const R3 domega = mk3(angvel_j.x - angvel_i.x, angvel_j.y - angvel_i.y,
                              angvel_j.z - angvel_i.z) * dt;
        const Real domega_n = dot(domega, mn);
        const Real omega_n = omega0_n + domega_n;
        const R3 domega_t = domega - domega_n * mn;
        const R3 omega_t = omega0_t + domega_t;
        const R3 t_n = -kt * J * omega_n * mn;
        const R3 t_t = -kn * I * omega_t;
        const R3 t_j = t_n + t_t;
        bool breaks;
        if (s->bond_break_ext > 0.0) {
            // The engine's option: normal extension (delta_n < 0 when the
            // pair separates) of bond_break_ext times the smaller radius.
            breaks = -delta_n >= s->bond_break_ext * fmin(x_i.w, x_j.w);
        } else {
            const Real sigma_max = length(f_n) / A + length(t_t) * R_bar / I;
            const Real tau_max = length(f_t) / A + length(t_n) * R_bar / J;
            breaks = sigma_max >= P.sigma_b || tau_max >= P.tau_b;
        }
        if (breaks) {
            s->bonds[2 * b] = P.nb0;  // src/cohesion.cuh:205 (sets x to nb0)
            continue;
        }
        s->bonds_delta[b] = mk4(delta_t.x, delta_t.y, delta_t.z, delta_n);
        s->bonds_omega[b] = mk4(omega_t.x, omega_t.y, omega_t.z, omega_n);
        R4& Fi = s->force[bx];
        Fi.x += f.x; Fi.y += f.y; Fi.z += f.z; Fi.w += 0.0;
        R4& Fj = s->force[by];
        Fj.x -= f.x; Fj.y -= f.y; Fj.z -= f.z; Fj.w -= 0.0;
        R4& Ti = s->torque[bx];
        Ti.x -= t_j.x; Ti.y -= t_j.y; Ti.z -= t_j.z; Ti.w -= 0.0;
        R4& Tj = s->torque[by];
        Tj.x += t_j.x; Tj.y += t_j.y; Tj.z += t_j.z; Tj.w += 0.0;
    }
}

// ---------------------------------------------------------------------------
// Wall reductions

// summation, src/integration.cuh:309-341, launched <<<blocks, 256>>> over np
// elements (src/device.cu:2445,2504): one element per thread, then the
// shared-memory halving tree.
void block_tree_sum(const Real* in, uint32_t np, std::vector<Real>& out) {
    const uint32_t nb = (np + 255u) / 256u;
    out.assign(nb, 0.0);
    Real cache[256];
    for (uint32_t b = 0; b < nb; ++b) {
        for (uint32_t t = 0; t < 256; ++t) {
            Real temp = 0.0;
            const uint32_t idx = b * 256u + t;
            if (idx < np) temp += in[idx];
            cache[t] = temp;
        }
        for (uint32_t i = 128; i != 0; i /= 2)
            for (uint32_t t = 0; t < i; ++t) cache[t] += cache[t + i];
        out[b] = cache[0];
    }
}

// findShearStressOnFixedMovingParticles, src/integration.cuh:457-484.
void stage_shear_stress(orc_sim* s) {
    const uint32_t np = s->np;
    for (uint32_t i = 0; i < np; ++i) {
        const Real z = s->x[i].z, fixvel = s->vel[i].w, force_x = s->force[i].x;
        Real f_x = 0.0;
        if (fixvel > 0.0 && z > s->L[2] * 0.5) f_x = force_x;
        s->tau_pp[i] = f_x / (s->L[0] * s->L[1]);
    }
    block_tree_sum(s->tau_pp.data(), np, s->tau_partial);
}

// integrate (TY3), src/integration.cuh:14-305.
void stage_integrate_particles(orc_sim* s) {
    const int64_t np = s->np;
    const Params& P = s->P;
    const int wall0mode = s->nw > 0 ? s->wmode[0] : 0;  // P11
    const Real wall0mass = s->nw > 0 ? s->w_mvfd[0].x : 0.0;
    Real tau_eff_x = 0.0;
    if (s->nw > 0 && wall0mode == 3)
        for (uint32_t i = 0; i < s->blocks(); ++i) tau_eff_x += s->tau_partial[i];
    const Real dt = s->dt;
    const R3 origo = mk3(s->origo[0], s->origo[1], s->origo[2]);
    const R3 L = mk3(s->L[0], s->L[1], s->L[2]);
    const int cvs = s->change_velocity_state;
    const Real velocity_factor = 10.0;
    const int64_t iter = s->iter;
#pragma omp parallel for schedule(static) if (np > 4096)
    for (int64_t idx = 0; idx < np; ++idx) {
        const uint32_t orig = s->perm[idx];
        const R4 force = s->force[orig];
        const R4 torque = s->torque[orig];
        const R4 angpos = s->angpos[orig];
        const R4 x = s->x_s[idx];
        const R4 vel = s->vel_s[idx];
        const R4 angvel = s->angvel_s[idx];
        R4 xyzsum = s->xyzsum[orig];
        R4 acc0, angacc0;
        if (iter == 0) {
            acc0 = mk4(0.0, 0.0, 0.0, 0.0);
            angacc0 = mk4(0.0, 0.0, 0.0, 0.0);
        } else {
            acc0 = s->acc[orig];
            angacc0 = s->angacc[orig];
        }
        const Real radius = x.w;
        R4 x_new, vel_new, angpos_new, angvel_new;
        const Real m = 4.0 / 3.0 * PI * radius * radius * radius * P.rho;
        R4 acc = mk4(0.0, 0.0, 0.0, 0.0);
        acc.x = force.x / m + P.g[0];
        acc.y = force.y / m + P.g[1];
        acc.z = force.z / m + P.g[2];
        R4 angacc = mk4(0.0, 0.0, 0.0, 0.0);
        angacc.x = torque.x * 1.0 / (2.0 / 5.0 * m * radius * radius);
        angacc.y = torque.y * 1.0 / (2.0 / 5.0 * m * radius * radius);
        angacc.z = torque.z * 1.0 / (2.0 / 5.0 * m * radius * radius);
        if (wall0mode == 3 && vel.w > 0.0001 && x.z > s->L[2] * 0.5) {
            const Real f_tau_x = (s->tau_x + tau_eff_x) * s->L[0] * s->L[1];
            acc.x = f_tau_x / wall0mass;
            acc.y = 0.0;
            acc.z -= P.g[2];
            angacc = mk4(0.0, 0.0, 0.0, 0.0);
        } else if (vel.w > 0.0001 && vel.w < 10.0) {
            acc.x = 0.0;
            acc.y = 0.0;
            acc.z -= P.g[2];
            angacc = mk4(0.0, 0.0, 0.0, 0.0);
        }
        if (vel.w >= 10.0) {
            acc.y = 0.0;
            angacc.x = 0.0;
            angacc.z = 0.0;
        }
        if (vel.w < -0.0001) {
            acc.x = 0.0; acc.y = 0.0; acc.z = 0.0;
            angacc = mk4(0.0, 0.0, 0.0, 0.0);
        }
        const R3 dacc_dt = mk3((acc.x - acc0.x) / dt, (acc.y - acc0.y) / dt, (acc.z - acc0.z) / dt);
        const R3 dangacc_dt = mk3((angacc.x - angacc0.x) / dt, (angacc.y - angacc0.y) / dt,
                                  (angacc.z - angacc0.z) / dt);
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
        if (s->periodic == 1) {
            if (x_new.x < origo.x) x_new.x += L.x;
            if (x_new.x > L.x) x_new.x -= L.x;
            if (x_new.y < origo.y) x_new.y += L.y;
            if (x_new.y > L.y) x_new.y -= L.y;
        } else if (s->periodic == 2) {
            if (x_new.x < origo.x) x_new.x += L.x;
            if (x_new.x > L.x) x_new.x -= L.x;
        }
        if (vel.w > 5.0 && vel.w < 10.0) {
            if (cvs == 1) vel_new.x *= velocity_factor;
            else if (cvs == -1) vel_new.x /= velocity_factor;
        }
        s->xyzsum[orig] = xyzsum;
        s->acc[orig] = acc;
        s->angacc[orig] = angacc;
        s->angvel[orig] = angvel_new;
        s->vel[orig] = vel_new;
        s->angpos[orig] = angpos_new;
        s->x[orig] = x_new;
    }
}

// Host loop around integrate: wmode-3 shear stress (src/device.cu:2427-2452)
// and the rate-state switch (src/device.cu:2454-2462,2486-2487).
void stage_integrate(orc_sim* s) {
    if (s->nw > 0 && s->wmode[0] == 3) stage_shear_stress(s);
    const Real v2_start = 10.0, v2_end = 15.0;
    if (s->t >= v2_start && s->t < v2_end && s->velocity_state == 1) {
        s->change_velocity_state = 1;
        s->velocity_state = 2;
    } else if (s->t >= v2_end && s->velocity_state == 2) {
        s->change_velocity_state = -1;
        s->velocity_state = 1;
    }
    stage_integrate_particles(s);
    s->change_velocity_state = 0;
}

// summation + integrateWalls, src/integration.cuh:309-450 and
// src/device.cu:2500-2527.  Every dynamic wall integrates wall 0's sum (P8);
// with g_wall_sums_per_wall (the engine's optional correction, no reference
// counterpart) wall w integrates the same tree-and-serial sum of its own
// per-particle forces.
void stage_walls(orc_sim* s) {
    if (s->nw == 0) return;
    block_tree_sum(s->force_pp.data(), s->np, s->force_partial);
    const Params& P = s->P;
    const Real dt = s->dt;
    std::vector<Real> own;
    for (uint32_t w = 0; w < s->nw; ++w) {
        R4 nx = s->w_nx[w], mvfd = s->w_mvfd[w];
        const int wmode = s->wmode[w];
        if (wmode == 0) continue;
        Real acc0;
        if (s->iter == 0) acc0 = 0.0; else acc0 = s->walls_acc[w];
        const Real* part = s->force_partial.data();
        if (g_wall_sums_per_wall && w > 0) {
            block_tree_sum(s->force_pp.data() + (size_t)w * s->np, s->np, own);
            part = own.data();
        }
        mvfd.z = 0.0;
        for (uint32_t i = 0; i < s->blocks(); ++i) mvfd.z += part[i];
        const Real sigma_0 = mvfd.w + P.sigma0_A * sin(2.0 * PI * P.sigma0_f * s->t);
        const Real N = -sigma_0 * s->L[0] * s->L[1];
        Real acc = (mvfd.z + N) / mvfd.x;
        if (wmode == 2) acc = 0.0;
        const Real dacc_dt = (acc - acc0) / dt;
        nx.w += mvfd.y * dt + 0.5 * acc * dt * dt + 1.0 / 6.0 * dacc_dt * dt * dt * dt;
        mvfd.y += acc * dt + 0.5 * dacc_dt * dt * dt;
        s->w_nx[w] = nx;
        s->w_mvfd[w] = mvfd;
        s->walls_acc[w] = acc;
    }
}

// DEM::updateGridSize, src/device.cu:363-393 (adaptive grid, nw > 0).
void stage_adaptive(orc_sim* s) {
    if (s->adaptive == 1 && s->nw > 0) s->L[2] = val(s->w_nx[0].w);
}

void advance(orc_sim* s) {
    s->t += s->dt;
    ++s->iter;
}

void one_step(orc_sim* s) {
    if (s->np > 0) {
        stage_adaptive(s);
        stage_hash(s);
        stage_sort(s);
        stage_reorder(s);
        if (s->P.contactmodel == 2 || s->P.contactmodel == 3) stage_topology(s);
        stage_interact(s);
        if (s->P.nb0 > 0) stage_bonds(s);
        stage_integrate(s);
        stage_walls(s);
    }
    advance(s);
}

// Field access ---------------------------------------------------------------

int64_t field_bytes(const orc_sim* s, int f) {
    const int64_t np = s->np, nw = s->nw, nb = s->P.nb0;
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
        case SPH_F_CELLSTART: case SPH_F_CELLEND: return (int64_t)s->ncells() * 4;
        case SPH_F_WALLS_NX: case SPH_F_WALLS_MVFD: return nw * 32;
        case SPH_F_WALLS_ACC: return nw * 8;
        case SPH_F_WALLS_FORCE_PP: return nw * np * 8;
        case SPH_F_BONDS: return nb * 8;
        case SPH_F_BONDS_DELTA: case SPH_F_BONDS_OMEGA: return nb * 32;
        case SPH_F_COLOR: return np * 4;
        default: return -1;
    }
}

template <class T> void copy_out(const std::vector<T>& v, void* dst, size_t n) {
    if (n) memcpy(dst, v.data(), n);
}
template <class T> void copy_in(std::vector<T>& v, const void* src, size_t n) {
    if (n) memcpy(v.data(), src, n);
}
#ifdef ORACLE_DUAL
// The tangent build stores Duals: fields cross the C ABI as their values.
void copy_out(const std::vector<Real>& v, void* dst, size_t) {
    double* d = static_cast<double*>(dst);
    for (size_t i = 0; i < v.size(); ++i) d[i] = v[i].v;
}
void copy_out(const std::vector<R4>& v, void* dst, size_t) {
    double* d = static_cast<double*>(dst);
    for (size_t i = 0; i < v.size(); ++i) {
        d[4 * i] = v[i].x.v; d[4 * i + 1] = v[i].y.v; d[4 * i + 2] = v[i].z.v; d[4 * i + 3] = v[i].w.v;
    }
}
void copy_in(std::vector<Real>& v, const void* src, size_t) {
    const double* d = static_cast<const double*>(src);
    for (size_t i = 0; i < v.size(); ++i) v[i] = d[i];
}
void copy_in(std::vector<R4>& v, const void* src, size_t) {
    const double* d = static_cast<const double*>(src);
    for (size_t i = 0; i < v.size(); ++i) v[i] = mk4(d[4 * i], d[4 * i + 1], d[4 * i + 2], d[4 * i + 3]);
}
#endif
void copy3_out(const std::vector<R4>& v, void* dst) {
    double* d = static_cast<double*>(dst);
    for (size_t i = 0; i < v.size(); ++i) { d[3 * i] = val(v[i].x); d[3 * i + 1] = val(v[i].y); d[3 * i + 2] = val(v[i].z); }
}
void copy3_in(std::vector<R4>& v, const void* src) {
    const double* d = static_cast<const double*>(src);
    for (size_t i = 0; i < v.size(); ++i) { v[i].x = d[3 * i]; v[i].y = d[3 * i + 1]; v[i].z = d[3 * i + 2]; }
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

orc_sim* orc_open_mem(const void* bytes, size_t nbytes, char* err, size_t errlen) {
    orc_sim* s = new orc_sim();
    std::string e;
    if (!parse_bin(s, bytes, nbytes, e)) {
        if (err && errlen) snprintf(err, errlen, "%s", e.c_str());
        delete s;
        return NULL;
    }
    init_runtime(s);
    return s;
}

void orc_close(orc_sim* s) { delete s; }

int orc_step(orc_sim* s, int64_t nsteps) {
#ifdef _OPENMP
    if (g_threads > 0) omp_set_num_threads(g_threads);
#endif
    for (int64_t k = 0; k < nsteps; ++k) one_step(s);
    return 0;
}

int orc_stage(orc_sim* s, int stage) {
    switch (stage) {
        case ORC_HASH: stage_adaptive(s); stage_hash(s); return 0;
        case ORC_SORT: stage_sort(s); return 0;
        case ORC_REORDER: stage_reorder(s); return 0;
        case ORC_TOPOLOGY:
            if (s->P.contactmodel == 2 || s->P.contactmodel == 3) stage_topology(s);
            return 0;
        case ORC_INTERACT: stage_interact(s); return 0;
        case ORC_BONDS: if (s->P.nb0 > 0) stage_bonds(s); return 0;
        case ORC_INTEGRATE: stage_integrate(s); return 0;
        case ORC_WALLS: stage_walls(s); return 0;
        case ORC_ADVANCE: advance(s); return 0;
        default: return SPH_E_ARG;
    }
}

int64_t orc_field_bytes(const orc_sim* s, int field) { return field_bytes(s, field); }

int orc_get_field(const orc_sim* s, int f, void* dst, size_t bytes) {
    const int64_t need = field_bytes(s, f);
    if (need < 0 || (size_t)need != bytes) return SPH_E_ARG;
    switch (f) {
        case SPH_F_X: copy_out(s->x, dst, bytes); return 0;
        case SPH_F_VEL: copy_out(s->vel, dst, bytes); return 0;
        case SPH_F_ANGVEL: copy3_out(s->angvel, dst); return 0;
        case SPH_F_ANGPOS: copy3_out(s->angpos, dst); return 0;
        case SPH_F_FORCE: copy3_out(s->force, dst); return 0;
        case SPH_F_TORQUE: copy3_out(s->torque, dst); return 0;
        case SPH_F_XYZSUM: copy3_out(s->xyzsum, dst); return 0;
        case SPH_F_ACC: copy3_out(s->acc, dst); return 0;
        case SPH_F_ANGACC: copy3_out(s->angacc, dst); return 0;
        case SPH_F_ES_DOT: copy_out(s->es_dot, dst, bytes); return 0;
        case SPH_F_ES: copy_out(s->es, dst, bytes); return 0;
        case SPH_F_EV_DOT: copy_out(s->ev_dot, dst, bytes); return 0;
        case SPH_F_EV: copy_out(s->ev, dst, bytes); return 0;
        case SPH_F_P: copy_out(s->p, dst, bytes); return 0;
        case SPH_F_CONTACTS: copy_out(s->contacts, dst, bytes); return 0;
        case SPH_F_DELTA_T: copy3_out(s->delta_t, dst); return 0;
        case SPH_F_DISTMOD: copy_out(s->distmod, dst, bytes); return 0;
        case SPH_F_CELLID: copy_out(s->cellid, dst, bytes); return 0;
        case SPH_F_PERM: copy_out(s->perm, dst, bytes); return 0;
        case SPH_F_CELLSTART: copy_out(s->cellStart, dst, bytes); return 0;
        case SPH_F_CELLEND: copy_out(s->cellEnd, dst, bytes); return 0;
        case SPH_F_WALLS_NX: copy_out(s->w_nx, dst, bytes); return 0;
        case SPH_F_WALLS_MVFD: copy_out(s->w_mvfd, dst, bytes); return 0;
        case SPH_F_WALLS_ACC: copy_out(s->walls_acc, dst, bytes); return 0;
        case SPH_F_WALLS_FORCE_PP: copy_out(s->force_pp, dst, bytes); return 0;
        case SPH_F_BONDS: copy_out(s->bonds, dst, bytes); return 0;
        case SPH_F_BONDS_DELTA: copy_out(s->bonds_delta, dst, bytes); return 0;
        case SPH_F_BONDS_OMEGA: copy_out(s->bonds_omega, dst, bytes); return 0;
        case SPH_F_COLOR: copy_out(s->color, dst, bytes); return 0;
        default: return SPH_E_ARG;
    }
}

int orc_set_field(orc_sim* s, int f, const void* src, size_t bytes) {
    const int64_t need = field_bytes(s, f);
    if (need < 0 || (size_t)need != bytes) return SPH_E_ARG;
    switch (f) {
        case SPH_F_X: copy_in(s->x, src, bytes); return 0;
        case SPH_F_VEL: copy_in(s->vel, src, bytes); return 0;
        case SPH_F_ANGVEL: copy3_in(s->angvel, src); return 0;
        case SPH_F_ANGPOS: copy3_in(s->angpos, src); return 0;
        case SPH_F_FORCE: copy3_in(s->force, src); return 0;
        case SPH_F_TORQUE: copy3_in(s->torque, src); return 0;
        case SPH_F_XYZSUM: copy3_in(s->xyzsum, src); return 0;
        case SPH_F_ACC: copy3_in(s->acc, src); return 0;
        case SPH_F_ANGACC: copy3_in(s->angacc, src); return 0;
        case SPH_F_ES_DOT: copy_in(s->es_dot, src, bytes); return 0;
        case SPH_F_ES: copy_in(s->es, src, bytes); return 0;
        case SPH_F_EV_DOT: copy_in(s->ev_dot, src, bytes); return 0;
        case SPH_F_EV: copy_in(s->ev, src, bytes); return 0;
        case SPH_F_P: copy_in(s->p, src, bytes); return 0;
        case SPH_F_CONTACTS: copy_in(s->contacts, src, bytes); return 0;
        case SPH_F_DELTA_T: copy3_in(s->delta_t, src); return 0;
        case SPH_F_DISTMOD: copy_in(s->distmod, src, bytes); return 0;
        case SPH_F_WALLS_NX: copy_in(s->w_nx, src, bytes); return 0;
        case SPH_F_WALLS_MVFD: copy_in(s->w_mvfd, src, bytes); return 0;
        case SPH_F_WALLS_ACC: copy_in(s->walls_acc, src, bytes); return 0;
        case SPH_F_BONDS: copy_in(s->bonds, src, bytes); return 0;
        case SPH_F_BONDS_DELTA: copy_in(s->bonds_delta, src, bytes); return 0;
        case SPH_F_BONDS_OMEGA: copy_in(s->bonds_omega, src, bytes); return 0;
        case SPH_F_COLOR: copy_in(s->color, src, bytes); return 0;
        default: return SPH_E_ARG;
    }
}

int64_t orc_write_mem(const orc_sim* s, void* dst, size_t cap) {
    Writer w;
    write_bin(s, w);
    if (dst && cap >= w.b.size()) memcpy(dst, w.b.data(), w.b.size());
    return (int64_t)w.b.size();
}

double orc_sqrt32(double v) { return val(sqrt32(v)); }

#ifdef ORACLE_DUAL
// Tangent build: seed a direction, step, read the directional derivatives.
// Parameters in the order k_n, k_t, gamma_n, gamma_t, mu_s, mu_d, gamma_wn,
// gamma_wt, mu_wd (the engine's adjoint parameters).
int orc_dual_seed_param(orc_sim* s, int which, double d) {
    Real* p[9] = {&s->P.k_n, &s->P.k_t, &s->P.gamma_n, &s->P.gamma_t, &s->P.mu_s,
                  &s->P.mu_d, &s->P.gamma_wn, &s->P.gamma_wt, &s->P.mu_wd};
    if (which < 0 || which > 8) return SPH_E_ARG;
    p[which]->d = d;
    return 0;
}
// field: SPH_F_X, SPH_F_VEL or SPH_F_ANGVEL; dir: np x 3.
int orc_dual_seed_field(orc_sim* s, int field, const double* dir) {
    std::vector<R4>* v = field == SPH_F_X ? &s->x : field == SPH_F_VEL ? &s->vel
                       : field == SPH_F_ANGVEL ? &s->angvel : nullptr;
    if (!v) return SPH_E_ARG;
    for (size_t i = 0; i < v->size(); ++i) {
        (*v)[i].x.d = dir[3 * i];
        (*v)[i].y.d = dir[3 * i + 1];
        (*v)[i].z.d = dir[3 * i + 2];
    }
    return 0;
}
int orc_dual_get_tangent(const orc_sim* s, int field, double* out) {
    const std::vector<R4>* v = field == SPH_F_X ? &s->x : field == SPH_F_VEL ? &s->vel
                             : field == SPH_F_ANGVEL ? &s->angvel : nullptr;
    if (!v) return SPH_E_ARG;
    for (size_t i = 0; i < v->size(); ++i) {
        out[3 * i] = (*v)[i].x.d;
        out[3 * i + 1] = (*v)[i].y.d;
        out[3 * i + 2] = (*v)[i].z.d;
    }
    return 0;
}
#endif
void orc_set_exact_length(int on) { g_exact_length = on != 0; }
void orc_set_contact_log(int on) { g_contact_log = on != 0; }
void orc_set_bond_model(orc_sim* s, double bond_k, double bond_break_ext) {
    s->bond_k = bond_k;
    s->bond_break_ext = bond_break_ext;
}
int orc_contact_log_width(void) { return CLOG_W; }
int orc_get_contact_log(const orc_sim* s, double* dst, size_t n) {
    const size_t want = (size_t)s->np * NC * CLOG_W;
    if (n != want || s->clog.size() != want) return -1;
    memcpy(dst, s->clog.data(), want * sizeof(double));
    return 0;
}
void orc_set_wall_sums_per_wall(int on) { g_wall_sums_per_wall = on != 0; }

double orc_tree_sum(const double* v, uint32_t n) {
    std::vector<Real> in(v, v + n), partial;
    block_tree_sum(in.data(), n, partial);
    Real sum = 0.0;
    for (size_t i = 0; i < partial.size(); ++i) sum += partial[i];
    return val(sum);
}

uint32_t orc_np(const orc_sim* s) { return s->np; }
uint32_t orc_nc(void) { return NC; }
double orc_time(const orc_sim* s) { return s->t; }
int64_t orc_iter(const orc_sim* s) { return s->iter; }

void orc_set_threads(int n) {
    g_threads = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

// ---------------------------------------------------------------------------
// Output-side readers.

// DEM::findOverlaps, src/sphere.cpp:674-724: every pair i < j (index order)
// with len(x_j - x_i) - (r_i + r_j) < 0 (FP64 root, no periodic images),
// both particles with fixvel exactly 0.  Writes at most `cap` pairs; returns
// the total in *count.
void orc_find_overlaps(const orc_sim* s, uint32_t* pairs, double* delta, uint64_t cap,
                       uint64_t* count) {
    uint64_t n = 0;
    for (uint32_t i = 0; i < s->np; ++i)
        for (uint32_t j = i + 1; j < s->np; ++j) {
            const R4 xi = s->x[i], xj = s->x[j];
            const Real dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
            const Real len = sqrt(dx * dx + dy * dy + dz * dz);
            const Real dn = len - (xi.w + xj.w);
            if (dn < 0.0 && (s->vel[i].w == 0 && s->vel[j].w == 0)) {
                if (n < cap) {
                    pairs[2 * n] = i;
                    pairs[2 * n + 1] = j;
                    delta[n] = val(dn);
                }
                ++n;
            }
        }
    *count = n;
}

// DEM::porosity, src/sphere.cpp:486-592 (with sphericalCap, :468-471, and
// r_max, :474-484, which returns the last particle's radius).  Returns -1
// when the slices are thinner than that radius times two (the reference
// exits), else 0.
int orc_porosity(const orc_sim* s, int z_slices, double* z_pos, double* porosity) {
    const double pi = 3.14159265358979323846;
    auto cap = [&](double h, double r) { return pi * h * h / 3.0 * (3.0 * r - h); };
    const double top = s->nw > 0 ? val(s->w_nx[0].w) : s->L[2];
    const double h_slice = (top - s->origo[2]) / (double)z_slices;
    const double r_last = s->np > 0 ? val(s->x[s->np - 1].w) : 0.0;
    if (h_slice <= r_last * 2.0) return -1;
    const double V_slice = h_slice * s->L[0] * s->L[1];
    for (int iz = 0; iz < z_slices; ++iz) {
        double V_void = V_slice;
        const double lo = iz * h_slice;
        const double hi = lo + h_slice;
        for (uint32_t i = 0; i < s->np; ++i) {
            const double zc = val(s->x[i].z), r = val(s->x[i].w);
            const double zl = zc - r, zh = zc + r;
            const double V_sphere = 4.0 / 3.0 * pi * r * r * r;
            if (lo <= zl && zh <= hi) {
                V_void -= V_sphere;
            } else {
                if (lo >= zc && lo <= zh)
                    V_void -= cap(zh - lo, r);
                else if (lo < zc && lo > zl)
                    V_void -= V_sphere + cap(lo - zl, r);
                if (hi <= zc && hi >= zl)
                    V_void -= cap(hi - zl, r);
                else if (hi > zc && hi < zh)
                    V_void -= V_sphere + cap(zh - hi, r);
            }
        }
        z_pos[iz] = lo + 0.5 * h_slice;
        porosity[iz] = V_void / V_slice;
    }
    return 0;
}

}  // extern "C"
