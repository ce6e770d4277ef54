// Reverse mode (adjoint) of the DEM step for contact model 2, with bonds and
// walls, on a recorded forward pass, for every simulation of a context.  The
// reference has no counterpart.
//
// The forward pass records, per step, the state the step reads (positions,
// velocities, angular velocities, the tangential displacements, the walls,
// and each bond's inputs) and the contact rows after the interaction pass,
// which hold exactly the contacts that produced a force.  The reverse pass
// then runs the steps backwards with that contact list frozen and every
// branch (normal-force clamp, stick or slip, wall friction, bond breakage)
// taken as the forward took it, and propagates cotangents:
//
//   integrate (TY3)     x' = x + v dt + (2/3) a dt^2 - (1/6) a0 dt^2
//                       v' = v + (3/2) a dt - (1/2) a0 dt,  a = F/m + g
//   contactLinear       F, T and the new tangential displacement of each
//                       contact, as functions of both particles' state, the
//                       stored displacement and (k_n, k_t, gamma_n, gamma_t,
//                       mu_s, mu_d)
//   contactLinear_wall  F, T of each wall contact and the force it puts on
//                       the wall (k_n, gamma_wn, gamma_wt, mu_wd), as
//                       functions of the particle's and the wall's state
//   bondsLinear         the bond's force and torque and its new
//                       displacements and rotations (k_n, k_t, gamma_n,
//                       gamma_t)
//   integrateWalls      p' = p + u dt + (2/3) a dt^2 - (1/6) a0 dt^2,
//                       u' = u + (3/2) a dt - (1/2) a0 dt, with
//                       a = (sum of wall forces - sigma0 Lx Ly) / m for a
//                       stress-controlled wall (wmode 1) and a = 0 for a
//                       velocity-controlled one (wmode 2)
//
// Lengths are differentiated as exact roots (the forward rounds them to
// FP32, which is locally constant and has no useful derivative).  Sums
// that span particles (parameter cotangents, the walls' position and
// velocity cotangents) are formed per workgroup and then per simulation in
// a fixed order, so gradients are reproducible bit for bit.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adjoint.hpp"
#include "dem_math.hpp"

namespace sphx {

namespace {

// The parameter table is read-only for every kernel: read through the
// constant address space (scalar loads that stores do not invalidate).
typedef const __attribute__((address_space(4))) DevParams CParams;

constexpr double PI = 3.14159265358979;  // src/constants.h:14

__device__ __forceinline__ d3 zero3() { return mk3(0.0, 0.0, 0.0); }

// Cotangents of the parameters, in the order of AdjParam: one work-item's
// running sums, kept in LDS (stride 256 doubles between parameters) to
// spare registers.
struct PGrad {
    double* s;
    __device__ __forceinline__ void add(int k, double x) const { s[k * 256] += x; }
};

// Reverse of contactLinear (src/contactmodels.cuh:244-411, kappa = 0,
// E <= 0.001) for one contact of particle a with partner b.  Inputs are the
// forward inputs; gF, gT are the cotangents of a's force and torque and gdt
// that of the new tangential displacement.  Outputs are accumulated.
template <bool PARAMS>
__device__ void adj_contact_linear(CParams& p, d3 va, d3 wa, d3 vb, d3 wb, double ra,
                                   double rb, d3 x, d3 dt0, d3 gF, d3 gT, d3 gdt, d3& gx,
                                   d3& gva, d3& gwa, d3& gvb, d3& gwb, d3& gdt0, PGrad& gp) {
    // forward (same expressions and FP32 lengths, so the same branches)
    const double len = length(x);
    const double delta = len - (ra + rb);
    const d3 n = x / len;
    const d3 vl = va - vb;
    const double ca = ra + delta / 2.0, cb = rb + delta / 2.0;
    const d3 na = cross(n, wa), nb = cross(n, wb);
    const d3 vel = vl + ca * na + cb * nb;
    const double vel_n = -dot(vl, n);
    const double nv = dot(n, vel);
    const d3 vel_t = vel - n * nv;
    const double vtl = length(vel_t);
    const double nd0 = dot(n, dt0);
    const d3 dtp = dt0 - n * nd0;
    const double dtpl = length(dtp);
    const double cn_raw = -p.k_n * delta + p.gamma_n * vel_n;
    const double cn = fmax(0.0, cn_raw);
    // kappa = 0: the capillary term is a signed zero and |f_n - f_c| = |f_n|.
    const double L = length(cn * n);
    const bool active = dtpl > 0.0 || vtl > 0.0;
    d3 dtmp = zero3(), f_raw = zero3(), t = zero3(), f_t = zero3();
    double frl = 0.0, limit = 0.0;
    bool slip = false;
    if (active) {
        dtmp = dtp + vel_t * p.dt;
        f_raw = -p.k_t * dtmp + -p.gamma_t * vel_t;
        frl = length(f_raw);
        limit = p.mu_s * L;
        slip = frl > limit;
        if (slip) {
            t = f_raw / frl;
            f_t = limit * t;
        } else {
            f_t = f_raw;
        }
    }
    // reverse
    d3 gn = zero3();
    double gdelta = 0.0;
    d3 gft = gF, gfn = gF;
    const double cu = -(ra + delta * 0.5);
    const d3 u = cu * n;
    gft += cross(gT, u);
    const d3 gu = cross(f_t, gT);
    gn += cu * gu;
    gdelta += -0.5 * dot(n, gu);
    d3 gdtp = zero3(), gvel_t = zero3();
    double gL = 0.0;
    if (active) {
        d3 gdtmp = zero3(), gfraw = zero3();
        if (slip) {
            // dt_new = -1/k_t (mu_d L t + gamma_t vel_t), f_t = mu_s L t
            const d3 A = p.mu_d * L * t + p.gamma_t * vel_t;
            const d3 gA = gdt * (-1.0 / p.k_t);
            if constexpr (PARAMS) gp.add(ADJ_KT, dot(gdt, A) / (p.k_t * p.k_t));
            if constexpr (PARAMS) gp.add(ADJ_MUD, L * dot(t, gA));
            gL += p.mu_d * dot(t, gA);
            d3 gt = (p.mu_d * L) * gA;
            if constexpr (PARAMS) gp.add(ADJ_GT, dot(vel_t, gA));
            gvel_t += p.gamma_t * gA;
            const double glim = dot(t, gft);
            gt += limit * gft;
            if constexpr (PARAMS) gp.add(ADJ_MUS, L * glim);
            gL += p.mu_s * glim;
            gfraw = (gt - t * dot(t, gt)) / frl;
        } else {
            gdtmp = gdt;
            gfraw = gft;
        }
        // f_raw = -k_t dtmp - gamma_t vel_t, dtmp = dtp + vel_t dt
        gdtmp += -p.k_t * gfraw;
        if constexpr (PARAMS) gp.add(ADJ_KT, -dot(dtmp, gfraw));
        if constexpr (PARAMS) gp.add(ADJ_GT, -dot(vel_t, gfraw));
        gvel_t += -p.gamma_t * gfraw;
        gdtp += gdtmp;
        gvel_t += p.dt * gdtmp;
    } else {
        gdtp += gdt;
    }
    if (L > 0.0) gfn += (gL / L) * (cn * n);
    const double gcn = dot(n, gfn);
    gn += cn * gfn;
    double gvel_n = 0.0;
    if (cn_raw > 0.0) {
        if constexpr (PARAMS) gp.add(ADJ_KN, -delta * gcn);
        gdelta += -p.k_n * gcn;
        if constexpr (PARAMS) gp.add(ADJ_GN, vel_n * gcn);
        gvel_n = p.gamma_n * gcn;
    }
    // dtp = dt0 - n (n . dt0)
    gdt0 += gdtp - n * dot(n, gdtp);
    gn += -(nd0 * gdtp + dot(n, gdtp) * dt0);
    // vel_t = vel - n (n . vel)
    const d3 gvel = gvel_t - n * dot(n, gvel_t);
    gn += -(nv * gvel_t + dot(n, gvel_t) * vel);
    // vel_n = -(vl . n)
    d3 gvl = -gvel_n * n;
    gn += -gvel_n * vl;
    // vel = vl + ca (n x wa) + cb (n x wb)
    gvl += gvel;
    gdelta += 0.5 * dot(gvel, na) + 0.5 * dot(gvel, nb);
    const d3 gna = ca * gvel, gnb = cb * gvel;
    gn += cross(wa, gna) + cross(wb, gnb);
    gwa += cross(gna, n);
    gwb += cross(gnb, n);
    gva += gvl;
    gvb += -gvl;
    // n = x / |x|, delta = |x| - (ra + rb)
    gx += (gn - n * dot(n, gn)) / len + n * gdelta;
}

// Reverse of contactLinear_wall (src/contactmodels.cuh:10-117) for one wall
// with normal n at distance delta from the particle (delta = e . x + const,
// with the wall position entering const), moving at speed wv, whose return
// value dot(f_n, n) (the force on the wall) has cotangent gret.  Adds to the
// particle's cotangents; returns the cotangents of delta and of wv.
__device__ void adj_contact_wall(CParams& p, d3 va, d3 wa, double ra, d3 n, double delta,
                                 double wv, d3 gF, d3 gT, double gret, d3& gva, d3& gwa,
                                 double& gdelta, double& gwv, PGrad& gp) {
    const d3 vl = va;
    const double c = ra + delta / 2.0;
    const d3 nw = cross(n, wa);
    const d3 vel = vl + c * nw + wv;
    const double vel_n = dot(vl, n);
    const double nv = dot(n, vel);
    const d3 vel_t = vel - n * nv;
    const double vtl = length(vel_t);
    const double cn_raw = -p.k_n * delta - p.gamma_wn * vel_n;
    const double cn = fmax(0.0, cn_raw);
    const d3 f_n = cn * n;
    const double fnl = length(f_n);
    d3 f_t = zero3();
    bool fric = vtl > 0.0 && p.gamma_wt > 0.0, stick = false;
    double limit = 0.0;
    if (fric) {
        const d3 f_t_visc = p.gamma_wt * vel_t;
        limit = p.mu_wd * fnl;
        stick = length(f_t_visc) < limit;
        f_t = stick ? -1.0 * f_t_visc : -limit * vel_t / vtl;
    }
    d3 gft = gF, gfn = gF + gret * n;
    gdelta = 0.0;
    const double cu = -(ra + delta * 0.5);
    gft += cross(gT, cu * n);
    gdelta += -0.5 * dot(n, cross(f_t, gT));
    d3 gvel_t = zero3();
    if (fric) {
        if (stick) {
            gvel_t += -p.gamma_wt * gft;
            gp.add(ADJ_GWT, -dot(vel_t, gft));
        } else {
            const d3 s = vel_t / vtl;
            const double glim = -dot(s, gft);
            const d3 gs = -limit * gft;
            gvel_t += (gs - s * dot(s, gs)) / vtl;
            gp.add(ADJ_MUWD, fnl * glim);
            if (fnl > 0.0) gfn += (p.mu_wd * glim / fnl) * f_n;
        }
    }
    const double gcn = dot(n, gfn);
    double gvel_n = 0.0;
    if (cn_raw > 0.0) {
        gp.add(ADJ_KN, -delta * gcn);
        gdelta += -p.k_n * gcn;
        gp.add(ADJ_GWN, -vel_n * gcn);
        gvel_n = -p.gamma_wn * gcn;
    }
    const d3 gvel = gvel_t - n * dot(n, gvel_t);
    d3 gvl = gvel_n * n + gvel;
    gdelta += 0.5 * dot(gvel, nw);
    gwa += cross(c * gvel, n);
    gva += gvl;
    // vel = vl + c (n x w) + wv: the wall speed is added to every component.
    gwv = gvel.x + gvel.y + gvel.z;
}

constexpr int ALPP = 4;               // work-items per particle
constexpr int ASPL = NC / ALPP;       // contact slots per work-item
static_assert(ADJ_PPB * ALPP == 256, "four work-items per particle");

__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int o = 1; o < ALPP; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ d3 group_sum(d3 v) {
    return mk3(group_sum(v.x), group_sum(v.y), group_sum(v.z));
}

// Force and torque cotangents of one particle (reverse of TY3 for the
// accelerations the step computes from them): from the cotangents of its
// new position, velocity and angular velocity and of the new accelerations
// (which also feed the next step as acc0).
__device__ __forceinline__ void force_cotangents(CParams& p, double r, double fv,
                                                 double4 Lx, double4 Lv, double4 Lw, double4 La,
                                                 double4 Laa, d3& gF, d3& gT) {
    const double dt = p.dt;
    const double m = 4.0 / 3.0 * PI * r * r * r * p.rho;
    const double I = 2.0 / 5.0 * m * r * r;
    d3 ga = mk3(La.x + (2.0 / 3.0) * dt * dt * Lx.x + 1.5 * dt * Lv.x,
                La.y + (2.0 / 3.0) * dt * dt * Lx.y + 1.5 * dt * Lv.y,
                La.z + (2.0 / 3.0) * dt * dt * Lx.z + 1.5 * dt * Lv.z);
    d3 gaa = mk3(Laa.x + 1.5 * dt * Lw.x, Laa.y + 1.5 * dt * Lw.y, Laa.z + 1.5 * dt * Lw.z);
    // fixvel modes (src/integration.cuh): the masked components do not
    // depend on the force.
    if (fv > 0.0001 && fv < 10.0) {
        ga.x = 0.0;
        ga.y = 0.0;
        gaa = zero3();
    }
    if (fv >= 10.0) {
        ga.y = 0.0;
        gaa.x = 0.0;
        gaa.z = 0.0;
    }
    if (fv < -0.0001) {
        ga = zero3();
        gaa = zero3();
    }
    gF = mk3(ga.x / m, ga.y / m, ga.z / m);
    gT = mk3(gaa.x / I, gaa.y / I, gaa.z / I);
}

__device__ __forceinline__ d3 image_shift(CParams& p, uint32_t wrd) {
    const uint32_t cx = (wrd >> CW_SHIFT_X) & 3u, cy = (wrd >> CW_SHIFT_Y) & 3u;
    return mk3(cx == 1u ? 0.0 + p.L[0] : cx == 2u ? 0.0 - p.L[0] : 0.0,
               cy == 1u ? 0.0 + p.L[1] : cy == 2u ? 0.0 - p.L[1] : 0.0, 0.0);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One list wall's contact (walls 0..nw-1 of the simulation): the particle's
// part, the wall's position and velocity cotangents (sign: how the contact
// distance changes with the wall position), the force sum's cotangent gS.
__device__ __forceinline__ void wall_contact(CParams& p, const AdjStep& st,
                                             const AdjState& g, int w, d3 va, d3 wa, double ra,
                                             d3 e, double sign, double dw, d3 wn, double wvel,
                                             d3 gF, d3 gT, d3& gxa, d3& gva, d3& gwa,
                                             double (&cwp)[MAXWALLS], double (&cwu)[MAXWALLS],
                                             PGrad& gp) {
    const bool listed = (uint32_t)w < p.nw;
    const double gS = listed ? g.wall[(p.wall_base + w) * ADJ_WALL + 3] : 0.0;
    double gdel = 0.0, gwv = 0.0;
    adj_contact_wall(p, va, wa, ra, wn, dw, wvel, gF, gT, gS, gva, gwa, gdel, gwv, gp);
    gxa += gdel * e;
    if (listed) {
#pragma unroll
        for (int k = 0; k < MAXWALLS; ++k)
            if (k == w) {
                cwp[k] += sign * gdel;
                cwu[k] += gwv;
            }
    }
}

// Reverse of one whole step, ALPP work-items per particle, in the step's
// sorted order (partner gathers then stay near each other in the recorded
// state).  Every particle's input cotangent is gathered by its own group,
// with no atomics:
//  1. integrate (TY3): from the cotangents of the step's outputs (Lout,
//     La/Laa out) the force and torque cotangents and the integrate part
//     of the input cotangents; the acceleration cotangents of the step's
//     inputs go to La/Laa in;
//  2. contacts: each work-item takes slots q, q + ALPP, ... of the row the
//     step left.  For contact a -> b it reverses a's contactLinear (a's
//     part, the parameters, and the stored displacement's cotangent, handed
//     to the step before where that slot held the same partner then), and
//     finds a in b's row to reverse b's contactLinear against a for the
//     part that falls on a (b's force cotangents recomputed from b's Lout).
//     Contact rows are symmetric unless a row overflowed; a contact whose
//     partner does not list it is counted in `asym` (the host then refuses
//     the result);
//  3. walls, on work-item 0, including the cotangent of each wall force
//     the particle puts on a list wall (from the wall's equation of
//     motion) and the cotangents of the walls' position and velocity;
//  4. the group's sums are written to Lin; the workgroup's partial sums of
//     the parameter and wall cotangents to part.
__global__ __launch_bounds__(256) void k_adj_step(AdjStep st, AdjState g) {
    __shared__ double red[ADJ_NPART][4];
    __shared__ double psum[ADJ_NPARAM * 256];
    CParams& p = *(CParams*)(st.ptab + st.chunk_sim[blockIdx.x / (256 / ADJ_PPB)]);
    const uint32_t q = threadIdx.x % ALPP;
    const uint32_t ia = blockIdx.x * ADJ_PPB + threadIdx.x / ALPP;
    PGrad gp;
    gp.s = psum + threadIdx.x;
    for (int k = 0; k < ADJ_NPARAM; ++k) gp.s[k * 256] = 0.0;
    double cwp[MAXWALLS], cwu[MAXWALLS];
#pragma unroll
    for (int k = 0; k < MAXWALLS; ++k) cwp[k] = cwu[k] = 0.0;
    if (ia - p.base < p.np) {
        const uint32_t a = st.perm[ia];
        const double4 xa4 = st.x[ia], va4 = st.v[ia];
        const d3 xa = xyz(xa4), va = xyz(va4), wa = xyz(st.w[ia]);
        const double ra = xa4.w;
        const double dt = p.dt;
        // 1. integrate
        const double4 Lx = g.Lout[0][a], Lv = g.Lout[1][a], Lw = g.Lout[2][a];
        d3 gF, gT;
        {
            const double4 La = g.La_out[a], Laa = g.Laa_out[a];
            force_cotangents(p, ra, va4.w, Lx, Lv, Lw, La, Laa, gF, gT);
        }
        d3 gxa = zero3(), gva = zero3(), gwa = zero3();
        if (q == 0) {
            // x' = x + ..., v' = v + ..., w' = w + ...
            gxa = mk3(Lx.x, Lx.y, Lx.z);
            gva = mk3(Lx.x * dt + Lv.x, Lx.y * dt + Lv.y, Lx.z * dt + Lv.z);
            gwa = mk3(Lw.x, Lw.y, Lw.z);
            const double c0 = st.first ? 0.0 : 1.0;  // acc0 is zero on a run's first step
            g.La_in[a] = make_double4(c0 * (-(dt * dt / 6.0) * Lx.x - 0.5 * dt * Lv.x),
                                      c0 * (-(dt * dt / 6.0) * Lx.y - 0.5 * dt * Lv.y),
                                      c0 * (-(dt * dt / 6.0) * Lx.z - 0.5 * dt * Lv.z), 0.0);
            g.Laa_in[a] = make_double4(c0 * -0.5 * dt * Lw.x, c0 * -0.5 * dt * Lw.y,
                                       c0 * -0.5 * dt * Lw.z, 0.0);
        }
        // 2. contacts.  This work-item's row words (and the previous step's)
        // are loaded at once; the occupied slots go into a small queue held
        // in registers (shifted, never indexed).
        // (The previous rows' presence is tested once, not per word: a
        // branch around each load would make the next load wait for it.)
        uint32_t wv[ASPL], pv[ASPL];
#pragma unroll
        for (int j = 0; j < ASPL; ++j) {
            wv[j] = st.rows[(size_t)a * NC + q + ALPP * j];
            pv[j] = EMPTY;
        }
        if (st.prev_rows) {
#pragma unroll
            for (int j = 0; j < ASPL; ++j) pv[j] = st.prev_rows[(size_t)a * NC + q + ALPP * j];
        }
        // Queue entry: the row word, and j | previous slot occupied << 3 |
        // same partner then << 4.
        uint32_t qw[ASPL], qj[ASPL];
        int nq = 0;
#pragma unroll
        for (int j = 0; j < ASPL; ++j) {
            qw[j] = EMPTY;
            qj[j] = 0;
        }
#pragma unroll
        for (int j = 0; j < ASPL; ++j) {
            const uint32_t wb = wv[j] & CW_IDX_MASK, pb = pv[j] & CW_IDX_MASK;
            const bool occ = wb != EMPTY;
            if (!occ && pb != EMPTY)
                g.Ldt_out[(size_t)a * NC + q + ALPP * j] = make_double4(0.0, 0.0, 0.0, 0.0);
            if (occ) {
#pragma unroll
                for (int k = ASPL - 1; k > 0; --k) {
                    qw[k] = qw[k - 1];
                    qj[k] = qj[k - 1];
                }
                qw[0] = wv[j];
                qj[0] = (uint32_t)j | (pb != EMPTY ? 8u : 0u) | (pb == wb ? 16u : 0u);
                ++nq;
            }
        }
#pragma unroll 1
        for (int k = 0; k < nq; ++k) {
            const uint32_t wrd = qw[0], fl = qj[0];
            const size_t mp = (size_t)a * NC + q + ALPP * (fl & 7u);
#pragma unroll
            for (int kk = 0; kk + 1 < ASPL; ++kk) {
                qw[kk] = qw[kk + 1];
                qj[kk] = qj[kk + 1];
            }
            const uint32_t b = wrd & CW_IDX_MASK;
            const uint32_t bs = st.srt[mp];  // the partner's sorted index
            const double4 xb4 = st.x[bs], vb4 = st.v[bs];
            const d3 xb = xyz(xb4), vb = xyz(vb4), wb = xyz(st.w[bs]);
            const double rb = xb4.w;
            // a's contact with b: a's part, parameters, displacement.
            {
                const d3 dm = image_shift(p, wrd);
                const d3 xab = mk3(xa.x - xb.x + dm.x, xa.y - xb.y + dm.y, xa.z - xb.z + dm.z);
                const d3 gdt = st.last ? zero3() : xyz(g.Ldt_in[mp]);
                d3 gx = zero3(), gvb = zero3(), gwb = zero3(), gdt0 = zero3();
                adj_contact_linear<true>(p, va, wa, vb, wb, ra, rb, xab, xyz(st.dt0[mp]), gF, gT,
                                         gdt, gx, gva, gwa, gvb, gwb, gdt0, gp);
                gxa += gx;
                if (fl & 8u)
                    g.Ldt_out[mp] = (fl & 16u) ? make_double4(gdt0.x, gdt0.y, gdt0.z, 0.0)
                                               : make_double4(0.0, 0.0, 0.0, 0.0);
            }
            // b's contact with a: the part that falls on a.
            uint32_t t = NC;
            const uint4* brow = reinterpret_cast<const uint4*>(st.rows + (size_t)b * NC);
            // Every word is read (no branch between the loads, which would
            // make each wait for the one before) and the first match kept.
#pragma unroll
            for (int h = NC / 4 - 1; h >= 0; --h) {
                const uint4 u = brow[h];
                t = (u.w & CW_IDX_MASK) == a ? 4 * h + 3 : t;
                t = (u.z & CW_IDX_MASK) == a ? 4 * h + 2 : t;
                t = (u.y & CW_IDX_MASK) == a ? 4 * h + 1 : t;
                t = (u.x & CW_IDX_MASK) == a ? 4 * h : t;
            }
            if (t == NC) {
                atomicAdd(g.asym, 1u);
                continue;
            }
            const size_t mpb = (size_t)b * NC + t;
            d3 gFb, gTb;
            force_cotangents(p, rb, vb4.w, g.Lout[0][b], g.Lout[1][b], g.Lout[2][b],
                             g.La_out[b], g.Laa_out[b], gFb, gTb);
            const d3 dmb = image_shift(p, st.rows[mpb]);
            const d3 xba = mk3(xb.x - xa.x + dmb.x, xb.y - xa.y + dmb.y, xb.z - xa.z + dmb.z);
            const d3 gdtb = st.last ? zero3() : xyz(g.Ldt_in[mpb]);
            d3 gx = zero3(), gvb_own = zero3(), gwb_own = zero3(), gva_part = zero3(),
               gwa_part = zero3(), gdt0 = zero3();
            adj_contact_linear<false>(p, vb, wb, va, wa, rb, ra, xba, xyz(st.dt0[mpb]), gFb, gTb,
                                      gdtb, gx, gvb_own, gwb_own, gva_part, gwa_part, gdt0, gp);
            gxa += -gx;
            gva += gva_part;
            gwa += gwa_part;
        }
        // 3. walls, as in the forward pass (src/contactsearch.cuh:602-701):
        // top (wall 0, or the implicit wall at L_z), bottom, then the side
        // walls by boundary mode; walls beyond nw are the implicit ones.
        if (q == 0) {
            const double4* wnx = st.wall_nx + p.wall_base;
            const double4* wmv = st.wall_mvfd + p.wall_base;
            {
                double wpos = p.L[2], wvel = 0.0;
                d3 wn = mk3(0.0, 0.0, -1.0);
                if (p.nw >= 1) { wn = xyz(wnx[0]); wpos = wnx[0].w; wvel = wmv[0].y; }
                const double dw = wpos - (xa.z + ra);
                if (dw < 0.0)
                    wall_contact(p, st, g, 0, va, wa, ra, mk3(0.0, 0.0, -1.0), 1.0, dw, wn, wvel,
                                 gF, gT, gxa, gva, gwa, cwp, cwu, gp);
            }
            {
                const double dw = xa.z - ra - p.origo[2];
                if (dw < 0.0)
                    wall_contact(p, st, g, -1, va, wa, ra, mk3(0.0, 0.0, 1.0), 0.0, dw,
                                 mk3(0.0, 0.0, 1.0), 0.0, gF, gT, gxa, gva, gwa, cwp, cwu, gp);
            }
            if (p.periodic == 0) {
                double wpos = p.L[0], wvel = 0.0;
                d3 wn = mk3(-1.0, 0.0, 0.0);
                if (p.nw >= 2) { wn = xyz(wnx[1]); wpos = wnx[1].w; wvel = wmv[1].y; }
                double dw = wpos - (xa.x + ra);
                if (dw < 0.0)
                    wall_contact(p, st, g, 1, va, wa, ra, mk3(-1.0, 0.0, 0.0), 1.0, dw, wn, wvel,
                                 gF, gT, gxa, gva, gwa, cwp, cwu, gp);
                wpos = 0.0; wvel = 0.0; wn = mk3(1.0, 0.0, 0.0);
                if (p.nw >= 3) { wn = xyz(wnx[2]); wpos = wnx[2].w; wvel = wmv[2].y; }
                dw = xa.x - ra - wpos;
                if (dw < 0.0)
                    wall_contact(p, st, g, 2, va, wa, ra, mk3(1.0, 0.0, 0.0), -1.0, dw, wn, wvel,
                                 gF, gT, gxa, gva, gwa, cwp, cwu, gp);
            }
            if (p.periodic == 0 || p.periodic == 2) {
                double wpos = p.L[1], wvel = 0.0;
                d3 wn = mk3(0.0, -1.0, 0.0);
                if (p.nw >= 4) { wn = xyz(wnx[3]); wpos = wnx[3].w; wvel = wmv[3].y; }
                double dw = wpos - (xa.y + ra);
                if (dw < 0.0)
                    wall_contact(p, st, g, 3, va, wa, ra, mk3(0.0, -1.0, 0.0), 1.0, dw, wn, wvel,
                                 gF, gT, gxa, gva, gwa, cwp, cwu, gp);
                wpos = 0.0; wvel = 0.0; wn = mk3(0.0, 1.0, 0.0);
                if (p.nw >= 5) { wn = xyz(wnx[4]); wpos = wnx[4].w; wvel = wmv[4].y; }
                dw = xa.y - ra - wpos;
                if (dw < 0.0)
                    wall_contact(p, st, g, 4, va, wa, ra, mk3(0.0, 1.0, 0.0), -1.0, dw, wn, wvel,
                                 gF, gT, gxa, gva, gwa, cwp, cwu, gp);
            }
        }
        // 4. the particle's input cotangents
        gxa = group_sum(gxa);
        gva = group_sum(gva);
        gwa = group_sum(gwa);
        if (q == 0) {
            g.Lin[0][a] = make_double4(gxa.x, gxa.y, gxa.z, 0.0);
            g.Lin[1][a] = make_double4(gva.x, gva.y, gva.z, 0.0);
            g.Lin[2][a] = make_double4(gwa.x, gwa.y, gwa.z, 0.0);
        }
    }
    // Partial sums of the workgroup: wavefront butterflies, then the four
    // wavefronts in order (a fixed order: reproducible bit for bit).
    const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    for (int k = 0; k < ADJ_NPARAM; ++k) {
        const double v = wave_sum(gp.s[k * 256]);
        if (lane == 0) red[k][wave] = v;
    }
#pragma unroll
    for (int k = 0; k < MAXWALLS; ++k) {
        const double vp = wave_sum(cwp[k]), vu = wave_sum(cwu[k]);
        if (lane == 0) {
            red[ADJ_NPARAM + 2 * k][wave] = vp;
            red[ADJ_NPARAM + 2 * k + 1][wave] = vu;
        }
    }
    __syncthreads();
    if (threadIdx.x < ADJ_NPART) {
        const int k = threadIdx.x;
        g.part[(size_t)blockIdx.x * ADJ_NPART + k] = red[k][0] + red[k][1] + red[k][2] + red[k][3];
    }
}

// Reverse of bondsLinear (src/cohesion.cuh:8-233) for the bonds that applied
// forces in the step, from the recorded inputs: the cotangents of the
// particles' forces and torques (recomputed from their output cotangents)
// and of the bond's new displacements and rotations.  Writes the
// cotangents of both particles' x, v, w to bcontrib, of the displacements
// and rotations it read to Lb_in, and its parameter cotangents to bpart.
// A bond that applied no force (broken, or broken in this step) leaves its
// displacements and rotations as they were: their cotangents pass through.
__global__ __launch_bounds__(256) void k_adj_bonds(AdjStep st, AdjState g) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b >= st.nbond) return;
    CParams& p = *(CParams*)(st.ptab + st.bond_sim[b]);
    const uint4 idx = st.bond_idx[b];
    double4* co = g.bcontrib + (size_t)6 * b;
    double* bp = g.bpart + (size_t)ADJ_NBPARAM * b;
    const double4 z4 = make_double4(0.0, 0.0, 0.0, 0.0);
    const double4 Ld = st.last ? z4 : g.Lb_out[2 * b], Lo = st.last ? z4 : g.Lb_out[2 * b + 1];
    if (!idx.z) {
        g.Lb_in[2 * b] = Ld;
        g.Lb_in[2 * b + 1] = Lo;
        for (int k = 0; k < 6; ++k) co[k] = z4;
        for (int k = 0; k < ADJ_NBPARAM; ++k) bp[k] = 0.0;
        return;
    }
    const uint32_t bi = idx.x, bj = idx.y;
    const double4* in = st.bond_in + (size_t)8 * b;
    const double4 x_i = in[0], x_j = in[1], v_i4 = in[2], v_j4 = in[3];
    const d3 w_i = xyz(in[4]), w_j = xyz(in[5]);
    const double4 d0 = in[6], o0 = in[7];
    const d3 v_i = xyz(v_i4), v_j = xyz(v_j4);
    const double dt = p.dt;
    // cotangents of the bond's force on i and torque t_j (force_i += f,
    // force_j -= f, torque_i -= t_j, torque_j += t_j)
    d3 gFi, gTi, gFj, gTj;
    force_cotangents(p, x_i.w, v_i4.w, g.Lout[0][bi], g.Lout[1][bi], g.Lout[2][bi],
                     g.La_out[bi], g.Laa_out[bi], gFi, gTi);
    force_cotangents(p, x_j.w, v_j4.w, g.Lout[0][bj], g.Lout[1][bj], g.Lout[2][bj],
                     g.La_out[bj], g.Laa_out[bj], gFj, gTj);
    const d3 gf = gFi - gFj, gtj = gTj - gTi;
    // forward
    const double R_bar = p.lambda_bar * fmin(x_i.w, x_j.w);
    const double A = PI * R_bar * R_bar;
    const double I = 0.25 * PI * R_bar * R_bar * R_bar * R_bar;
    const double J = I * 2.0;
    const d3 x = mk3(x_i.x - x_j.x, x_i.y - x_j.y, x_i.z - x_j.z);
    const double len = length(x);
    const double sep = len - (x_i.w + x_j.w);
    const double overlap = fmin(0.0, sep);
    const d3 n = x / len;
    const d3 dt0 = xyz(d0), om0 = xyz(o0);
    const double nd0 = dot(n, dt0);
    const d3 dt0p = dt0 - n * nd0;
    const double ci = x_i.w + overlap / 2.0, cj = x_j.w + overlap / 2.0;
    const d3 ni = cross(n, w_i), nj = cross(n, w_j);
    const d3 dd = ((v_i - v_j) + ci * ni + cj * nj) * dt;
    const double ddn = -dot(dd, n);
    const double dn = d0.w + ddn;
    const double ndd = dot(n, dd);
    const d3 ddt = dd - n * ndd;
    const d3 dtv = dt0p + ddt;
    // Stiffness per unit area (the bond option replaces k_n, k_t here; its
    // own cotangent is not formed, and then k_n, k_t get none from bonds).
    const bool ref_k = !(p.bond_k > 0.0);
    const double kn = ref_k ? p.k_n : p.bond_k / A;
    const double kt = ref_k ? p.k_t : p.bond_k / A;
    const double cfn = kn * A * dn + p.gamma_n * ddn / dt;
    const double no0 = dot(n, om0);
    const d3 dom = (w_j - w_i) * dt;
    const double ndo = dot(n, dom);
    const double don = -ndo;
    const double omn = o0.w + don;
    const d3 omt = (om0 - n * no0) + (dom - n * ndo);
    // reverse
    const d3 gdtv = xyz(Ld), gomt_o = xyz(Lo);
    double gdn = Ld.w, gomn = Lo.w;
    d3 gn = zero3();
    // t_j = t_n + t_t; t_n = k_t J omn n; t_t = -k_n I omt
    const d3 gomt = gomt_o + (-kn * I) * gtj;
    bp[0] = -I * dot(omt, gtj);                       // k_n
    const double ngt = dot(n, gtj);
    gomn += kt * J * ngt;
    bp[1] = J * omn * ngt;                            // k_t
    gn += (kt * J * omn) * gtj;
    // omt = om0 - n (n . om0) + dom - n (n . dom); omn = o0.w - n . dom
    d3 gom0 = gomt - n * dot(n, gomt);
    gn += -(no0 * gomt + dot(n, gomt) * om0);
    d3 gdom = gomt - n * dot(n, gomt);
    gn += -(ndo * gomt + dot(n, gomt) * dom);
    const double gom0n = gomn;
    gdom += -gomn * n;
    gn += -gomn * dom;
    d3 gwi = -dt * gdom, gwj = dt * gdom;
    // f = f_n + f_t; f_n = cfn n; f_t = -k_t A dtv - gamma_t ddt / dt
    const double gcfn = dot(n, gf);
    gn += cfn * gf;
    gdn += kn * A * gcfn;
    bp[0] += A * dn * gcfn;
    double gddn = p.gamma_n / dt * gcfn;
    bp[2] = ddn / dt * gcfn;                          // gamma_n
    const d3 gdtv_tot = gdtv + (-kt * A) * gf;
    bp[1] += -A * dot(dtv, gf);
    if (!ref_k) bp[0] = bp[1] = 0.0;
    d3 gddt = (-p.gamma_t / dt) * gf;
    bp[3] = -dot(ddt, gf) / dt;                       // gamma_t
    // dtv = dt0p + ddt; dn = d0.w + ddn
    const d3 gdt0p = gdtv_tot;
    gddt += gdtv_tot;
    const double gd0n = gdn;
    gddn += gdn;
    // ddt = dd - n (n . dd); ddn = -(dd . n)
    d3 gdd = gddt - n * dot(n, gddt);
    gn += -(ndd * gddt + dot(n, gddt) * dd);
    gdd += -gddn * n;
    gn += -gddn * dd;
    // dd = ((v_i - v_j) + ci (n x w_i) + cj (n x w_j)) dt
    const d3 gw = dt * gdd;
    const d3 gvi = gw, gvj = -gw;
    gn += cross(w_i, ci * gw) + cross(w_j, cj * gw);
    gwi += cross(ci * gw, n);
    gwj += cross(cj * gw, n);
    const double gov = 0.5 * (dot(gw, ni) + dot(gw, nj));
    // dt0p = dt0 - n (n . dt0)
    const d3 gdt0 = gdt0p - n * dot(n, gdt0p);
    gn += -(nd0 * gdt0p + dot(n, gdt0p) * dt0);
    // overlap = fmin(0, len - R); n = x / len
    const double glen = sep < 0.0 ? gov : 0.0;
    const d3 gx = (gn - n * dot(n, gn)) / len + n * glen;
    g.Lb_in[2 * b] = make_double4(gdt0.x, gdt0.y, gdt0.z, gd0n);
    g.Lb_in[2 * b + 1] = make_double4(gom0.x, gom0.y, gom0.z, gom0n);
    co[0] = make_double4(gx.x, gx.y, gx.z, 0.0);
    co[1] = make_double4(gvi.x, gvi.y, gvi.z, 0.0);
    co[2] = make_double4(gwi.x, gwi.y, gwi.z, 0.0);
    co[3] = make_double4(-gx.x, -gx.y, -gx.z, 0.0);
    co[4] = make_double4(gvj.x, gvj.y, gvj.z, 0.0);
    co[5] = make_double4(gwj.x, gwj.y, gwj.z, 0.0);
}

// Adds each particle's share of its bonds' cotangents to its input
// cotangents, in ascending bond id over the bonds the forward gather
// visited (the particle's own CSR list, or the whole list for particle nb0
// after a broken bond was retargeted onto it).
__global__ __launch_bounds__(256) void k_adj_bond_gather(AdjStep st, AdjState g) {
    CParams& p = *(CParams*)(st.ptab + st.chunk_sim[blockIdx.x]);
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i - p.base >= p.np) return;
    const bool full = (i - p.base == p.nb0) && st.bond_full[p.sim] != 0u;
    const uint32_t b0 = full ? p.bond_base : st.badj_off[i];
    const uint32_t b1 = full ? p.bond_base + p.nb0 : st.badj_off[i + 1];
    if (b0 == b1) return;
    double4 gx = g.Lin[0][i], gv = g.Lin[1][i], gw = g.Lin[2][i];
    for (uint32_t k = b0; k < b1; ++k) {
        const uint32_t b = full ? k : st.badj[k];
        const uint4 idx = st.bond_idx[b];
        if (!idx.z) continue;
        const double4* co = g.bcontrib + (size_t)6 * b;
        for (int side = 0; side < 2; ++side) {
            if ((side == 0 ? idx.x : idx.y) != i) continue;
            const double4 cx = co[3 * side], cv = co[3 * side + 1], cw = co[3 * side + 2];
            gx.x += cx.x; gx.y += cx.y; gx.z += cx.z;
            gv.x += cv.x; gv.y += cv.y; gv.z += cv.z;
            gw.x += cw.x; gw.y += cw.y; gw.z += cw.z;
        }
    }
    g.Lin[0][i] = gx;
    g.Lin[1][i] = gv;
    g.Lin[2][i] = gw;
}

// Fixed-order sum of n values v[k * stride] over one workgroup of 256: each
// work-item sums its strided share, then a tree.  Result on every work-item.
__device__ double block_sum(const double* v, uint32_t n, uint32_t stride, double* cache) {
    double t = 0.0;
    for (uint32_t k = threadIdx.x; k < n; k += 256) t += v[(size_t)k * stride];
    cache[threadIdx.x] = t;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) cache[threadIdx.x] += cache[threadIdx.x + h];
        __syncthreads();
    }
    const double r = cache[0];
    __syncthreads();
    return r;
}

// Per simulation (one workgroup each): the step's parameter cotangents
// (particle workgroups, then bonds) added to the running sums, and the
// reverse of integrateWalls (src/integration.cuh:344-450) for the walls:
// from the cotangents of the position, velocity and acceleration the step
// left (P, U, A) and of the wall contacts' position and velocity (cp, cu),
// those of the position, velocity and acceleration the step started from;
// then the force sum's cotangent for the step before (the wall's
// acceleration there is (sum + N) / m for wmode 1, zero for wmode 2, and
// wall w integrates wall 0's sum unless each wall has its own).
__global__ __launch_bounds__(256) void k_adj_reduce(AdjStep st, AdjState g) {
    __shared__ double cache[256];
    __shared__ double tot[ADJ_NPART];
    CParams& p = *(CParams*)(st.ptab + blockIdx.x);
    const uint32_t wb0 = p.base / ADJ_PPB, nwb = p.npad / ADJ_PPB;
    for (int k = 0; k < ADJ_NPART; ++k) {
        const double v = block_sum(g.part + (size_t)wb0 * ADJ_NPART + k, nwb, ADJ_NPART, cache);
        if (threadIdx.x == 0) tot[k] = v;
    }
    double bsum[ADJ_NBPARAM] = {0.0, 0.0, 0.0, 0.0};
    if (st.nbond && p.nb0)
        for (int k = 0; k < ADJ_NBPARAM; ++k)
            bsum[k] = block_sum(g.bpart + (size_t)p.bond_base * ADJ_NBPARAM + k, p.nb0,
                                ADJ_NBPARAM, cache);
    __syncthreads();
    if (threadIdx.x != 0) return;
    double* gp = g.gparam + (size_t)p.sim * ADJ_NPARAM;
    for (int k = 0; k < ADJ_NPARAM; ++k) gp[k] += tot[k] + (k < ADJ_NBPARAM ? bsum[k] : 0.0);
    const double dt = p.dt;
    double* W = g.wall + (size_t)p.wall_base * ADJ_WALL;
    for (uint32_t w = 0; w < p.nw; ++w) {
        double* s = W + w * ADJ_WALL;
        const double P = s[0], U = s[1], A = s[2];
        const double cp = tot[ADJ_NPARAM + 2 * w], cu = tot[ADJ_NPARAM + 2 * w + 1];
        if (p.wmode[w] == 0) {
            s[0] = P + cp;
            s[1] = U + cu;
        } else {
            s[0] = P + cp;
            s[1] = dt * P + U + cu;
            s[2] = st.first ? 0.0 : -(dt * dt / 6.0) * P - 0.5 * dt * U;
        }
        s[3] = 0.0;
    }
    for (uint32_t w = 0; w < p.nw; ++w) {
        if (p.wmode[w] != 1) continue;
        const double* s = W + w * ADJ_WALL;
        const double ga = s[2] + (2.0 / 3.0) * dt * dt * s[0] + 1.5 * dt * s[1];
        const uint32_t src = p.wall_sums ? w : 0u;
        W[src * ADJ_WALL + 3] += ga / st.wall_mvfd[p.wall_base + w].x;
    }
}

// Host rows <-> batch arrays.
__device__ __forceinline__ bool host_row(const AdjStep& st, uint32_t i, uint32_t& h) {
    CParams& p = *(CParams*)(st.ptab + st.chunk_sim[i / 256]);
    const uint32_t local = i - p.base;
    h = p.hbase + local;
    return local < p.np;
}

__global__ __launch_bounds__(256) void k_adj_expand(double4* dst, const double* src, AdjStep st) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= st.ntot) return;
    uint32_t h;
    dst[i] = host_row(st, i, h) ? make_double4(src[3 * h], src[3 * h + 1], src[3 * h + 2], 0.0)
                                : make_double4(0.0, 0.0, 0.0, 0.0);
}

__global__ __launch_bounds__(256) void k_adj_pack(double* dst, const double4* src, AdjStep st) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    uint32_t h;
    if (i >= st.ntot || !host_row(st, i, h)) return;
    const double4 v = src[i];
    dst[3 * h] = v.x;
    dst[3 * h + 1] = v.y;
    dst[3 * h + 2] = v.z;
}

// dst[i].xyz += the host row of particle i (a loss term on an intermediate
// state).
__global__ __launch_bounds__(256) void k_adj_accumulate(double4* dst, const double* src,
                                                        AdjStep st) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    uint32_t h;
    if (i >= st.ntot || !host_row(st, i, h)) return;
    double4 v = dst[i];
    v.x += src[3 * h];
    v.y += src[3 * h + 1];
    v.z += src[3 * h + 2];
    dst[i] = v;
}

}  // namespace

hipError_t launch_adj_accumulate(double4* dst, const double* src, const AdjStep& st,
                                 hipStream_t s) {
    hipLaunchKernelGGL(k_adj_accumulate, dim3((st.ntot + 255) / 256), dim3(256), 0, s, dst, src,
                       st);
    return hipGetLastError();
}

hipError_t launch_adj_expand(double4* dst, const double* src, const AdjStep& st, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_expand, dim3((st.ntot + 255) / 256), dim3(256), 0, s, dst, src, st);
    return hipGetLastError();
}

hipError_t launch_adj_pack(double* dst, const double4* src, const AdjStep& st, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_pack, dim3((st.ntot + 255) / 256), dim3(256), 0, s, dst, src, st);
    return hipGetLastError();
}

hipError_t launch_adj_step(const AdjStep& st, const AdjState& g, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_step, dim3(st.ntot / ADJ_PPB), dim3(256), 0, s, st, g);
    if (st.nbond) {
        hipLaunchKernelGGL(k_adj_bonds, dim3((st.nbond + 255) / 256), dim3(256), 0, s, st, g);
        hipLaunchKernelGGL(k_adj_bond_gather, dim3(st.ntot / 256), dim3(256), 0, s, st, g);
    }
    hipLaunchKernelGGL(k_adj_reduce, dim3(st.nsim), dim3(256), 0, s, st, g);
    return hipGetLastError();
}

}  // namespace sphx
