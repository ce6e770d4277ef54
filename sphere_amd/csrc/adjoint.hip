// Reverse mode (adjoint) of the DEM step for contact model 2, on a recorded
// forward pass.  The reference has no counterpart.
//
// The forward pass records, per step, the state the step reads (positions,
// velocities, angular velocities, the previous accelerations, the
// tangential displacements and the walls) and the contact rows after the
// interaction pass, which hold exactly the contacts that produced a force.
// The reverse pass then runs the steps backwards with that contact list
// frozen and every branch (normal-force clamp, stick or slip, wall
// friction) taken as the forward took it, and propagates cotangents:
//
//   integrate (TY3)     x' = x + v dt + (2/3) a dt^2 - (1/6) a0 dt^2
//                       v' = v + (3/2) a dt - (1/2) a0 dt,  a = F/m + g
//   contactLinear       F, T and the new tangential displacement of each
//                       contact, as functions of both particles' state, the
//                       stored displacement and (k_n, k_t, gamma_n, gamma_t,
//                       mu_s, mu_d)
//   contactLinear_wall  F, T of each wall contact (k_n, gamma_wn, gamma_wt,
//                       mu_wd)
//
// Lengths are differentiated as exact roots (the forward rounds them to
// FP32, which is locally constant and has no useful derivative).  Walls move
// along their recorded trajectories: the wall equation of motion is not
// differentiated, which is exact for walls at fixed position or velocity.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adjoint.hpp"
#include "dem_math.hpp"

namespace sphx {

namespace {

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
__device__ void adj_contact_linear(const DevParams& p, d3 va, d3 wa, d3 vb, d3 wb, double ra,
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
// with normal n, distance delta = e . x + const, wall speed wv.
__device__ void adj_contact_wall(const DevParams& p, d3 va, d3 wa, double ra, d3 n, d3 e,
                                 double delta, double wv, d3 gF, d3 gT, d3& gx, d3& gva,
                                 d3& gwa, PGrad& gp) {
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
    d3 gft = gF, gfn = gF;
    double gdelta = 0.0;
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
    gx += gdelta * e;
}

constexpr int ALPP = 4;               // work-items per particle
constexpr int ASPL = NC / ALPP;       // contact slots per work-item
constexpr int APPB = 256 / ALPP;      // particles per workgroup

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
__device__ __forceinline__ void force_cotangents(const DevParams& p, double r, double fv,
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

__device__ __forceinline__ d3 image_shift(const DevParams& p, uint32_t wrd) {
    const uint32_t cx = (wrd >> CW_SHIFT_X) & 3u, cy = (wrd >> CW_SHIFT_Y) & 3u;
    return mk3(cx == 1u ? 0.0 + p.L[0] : cx == 2u ? 0.0 - p.L[0] : 0.0,
               cy == 1u ? 0.0 + p.L[1] : cy == 2u ? 0.0 - p.L[1] : 0.0, 0.0);
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
//  3. walls, on work-item 0;
//  4. the group's sums are written to Lin.
__global__ __launch_bounds__(256) void k_adj_step(const DevParams p, AdjStep st, AdjState g,
                                                  double Lz) {
    __shared__ double red[ADJ_NPARAM][4];
    __shared__ double psum[ADJ_NPARAM * 256];
    const uint32_t q = threadIdx.x % ALPP;
    const uint32_t ia = blockIdx.x * APPB + threadIdx.x / ALPP;
    PGrad gp;
    gp.s = psum + threadIdx.x;
    for (int k = 0; k < ADJ_NPARAM; ++k) gp.s[k * 256] = 0.0;
    if (ia < p.np) {
        const uint32_t np = p.np;
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
        uint32_t wv[ASPL], pv[ASPL];
#pragma unroll
        for (int j = 0; j < ASPL; ++j) {
            const size_t mp = (size_t)a * NC + q + ALPP * j;
            wv[j] = st.rows[mp];
            pv[j] = st.prev_rows ? st.prev_rows[mp] : EMPTY;
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
            const bool occ = (wv[j] & CW_IDX_MASK) < np;
            if (!occ && (pv[j] & CW_IDX_MASK) < np)
                g.Ldt_out[(size_t)a * NC + q + ALPP * j] = make_double4(0.0, 0.0, 0.0, 0.0);
            if (occ) {
#pragma unroll
                for (int k = ASPL - 1; k > 0; --k) {
                    qw[k] = qw[k - 1];
                    qj[k] = qj[k - 1];
                }
                const uint32_t pb = pv[j] & CW_IDX_MASK;
                qw[0] = wv[j];
                qj[0] = (uint32_t)j | (pb < np ? 8u : 0u) |
                        (pb == (wv[j] & CW_IDX_MASK) ? 16u : 0u);
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
#pragma unroll
            for (int h = 0; h < NC / 4; ++h) {
                const uint4 u = brow[h];
                if (t == NC) {
                    if ((u.x & CW_IDX_MASK) == a) t = 4 * h;
                    else if ((u.y & CW_IDX_MASK) == a) t = 4 * h + 1;
                    else if ((u.z & CW_IDX_MASK) == a) t = 4 * h + 2;
                    else if ((u.w & CW_IDX_MASK) == a) t = 4 * h + 3;
                }
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
        // 3. walls, as in the forward pass: top (wall 0 or L_z), bottom,
        // then the side walls by boundary mode.
        if (q == 0) {
            {
                double wpos = Lz, wv = 0.0;
                d3 wn = mk3(0.0, 0.0, -1.0);
                if (p.nw >= 1) {
                    wn = xyz(st.wall_nx[0]);
                    wpos = st.wall_nx[0].w;
                    wv = st.wall_mvfd[0].y;
                }
                const double dw = wpos - (xa.z + ra);
                if (dw < 0.0)
                    adj_contact_wall(p, va, wa, ra, wn, mk3(0.0, 0.0, -1.0), dw, wv, gF, gT, gxa,
                                     gva, gwa, gp);
            }
            {
                const double dw = xa.z - ra - p.origo[2];
                if (dw < 0.0)
                    adj_contact_wall(p, va, wa, ra, mk3(0.0, 0.0, 1.0), mk3(0.0, 0.0, 1.0), dw,
                                     0.0, gF, gT, gxa, gva, gwa, gp);
            }
            if (p.periodic == 0) {
                double wpos = p.L[0], wv = 0.0;
                d3 wn = mk3(-1.0, 0.0, 0.0);
                if (p.nw >= 2) { wn = xyz(st.wall_nx[1]); wpos = st.wall_nx[1].w; wv = st.wall_mvfd[1].y; }
                double dw = wpos - (xa.x + ra);
                if (dw < 0.0)
                    adj_contact_wall(p, va, wa, ra, wn, mk3(-1.0, 0.0, 0.0), dw, wv, gF, gT, gxa,
                                     gva, gwa, gp);
                wpos = 0.0; wv = 0.0; wn = mk3(1.0, 0.0, 0.0);
                if (p.nw >= 3) { wn = xyz(st.wall_nx[2]); wpos = st.wall_nx[2].w; wv = st.wall_mvfd[2].y; }
                dw = xa.x - ra - wpos;
                if (dw < 0.0)
                    adj_contact_wall(p, va, wa, ra, wn, mk3(1.0, 0.0, 0.0), dw, wv, gF, gT, gxa,
                                     gva, gwa, gp);
            }
            if (p.periodic == 0 || p.periodic == 2) {
                double wpos = p.L[1], wv = 0.0;
                d3 wn = mk3(0.0, -1.0, 0.0);
                if (p.nw >= 4) { wn = xyz(st.wall_nx[3]); wpos = st.wall_nx[3].w; wv = st.wall_mvfd[3].y; }
                double dw = wpos - (xa.y + ra);
                if (dw < 0.0)
                    adj_contact_wall(p, va, wa, ra, wn, mk3(0.0, -1.0, 0.0), dw, wv, gF, gT, gxa,
                                     gva, gwa, gp);
                wpos = 0.0; wv = 0.0; wn = mk3(0.0, 1.0, 0.0);
                if (p.nw >= 5) { wn = xyz(st.wall_nx[4]); wpos = st.wall_nx[4].w; wv = st.wall_mvfd[4].y; }
                dw = xa.y - ra - wpos;
                if (dw < 0.0)
                    adj_contact_wall(p, va, wa, ra, wn, mk3(0.0, 1.0, 0.0), dw, wv, gF, gT, gxa,
                                     gva, gwa, gp);
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
    // Parameter cotangents: wavefront sums, then the four wavefronts' sums,
    // then one atomic per parameter and workgroup.
    const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    for (int k = 0; k < ADJ_NPARAM; ++k) {
        double v = gp.s[k * 256];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < ADJ_NPARAM) {
        const int k = threadIdx.x;
        atomicAdd(&g.gparam[k], red[k][0] + red[k][1] + red[k][2] + red[k][3]);
    }
}

// Three-component host layout <-> the engine's four-component one.
__global__ __launch_bounds__(256) void k_adj_expand(double4* dst, const double* src, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = make_double4(src[3 * i], src[3 * i + 1], src[3 * i + 2], 0.0);
}

__global__ __launch_bounds__(256) void k_adj_pack(double* dst, const double4* src, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const double4 v = src[i];
        dst[3 * i] = v.x;
        dst[3 * i + 1] = v.y;
        dst[3 * i + 2] = v.z;
    }
}

// dst[i].xyz += src[3 i .. 3 i + 2] (a loss term on an intermediate state).
__global__ __launch_bounds__(256) void k_adj_accumulate(double4* dst, const double* src,
                                                        uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        double4 v = dst[i];
        v.x += src[3 * i];
        v.y += src[3 * i + 1];
        v.z += src[3 * i + 2];
        dst[i] = v;
    }
}

}  // namespace

hipError_t launch_adj_accumulate(double4* dst, const double* src, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_accumulate, dim3((n + 255) / 256), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_adj_expand(double4* dst, const double* src, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_expand, dim3((n + 255) / 256), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_adj_pack(double* dst, const double4* src, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_pack, dim3((n + 255) / 256), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_adj_step(const DevParams& p, const AdjStep& st, const AdjState& g, double Lz,
                           hipStream_t s) {
    const dim3 grid((p.np + APPB - 1) / APPB), block(256);
    hipLaunchKernelGGL(k_adj_step, grid, block, 0, s, p, st, g, Lz);
    return hipGetLastError();
}

}  // namespace sphx
