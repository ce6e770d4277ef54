// Reverse-mode step (contact model 2, bonds, walls) over a recorded forward
// pass, for every simulation of a context.
#ifndef SPHERE_AMD_ADJOINT_HPP_
#define SPHERE_AMD_ADJOINT_HPP_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dem_types.hpp"

namespace sphx {

// Parameters with a cotangent, in output order (per simulation).
enum AdjParam : int {
    ADJ_KN = 0,    // k_n (contacts, walls and bonds)
    ADJ_KT = 1,    // k_t (contacts and bonds)
    ADJ_GN = 2,    // gamma_n
    ADJ_GT = 3,    // gamma_t
    ADJ_MUS = 4,   // mu_s
    ADJ_MUD = 5,   // mu_d
    ADJ_GWN = 6,   // gamma_wn
    ADJ_GWT = 7,   // gamma_wt
    ADJ_MUWD = 8,  // mu_wd
    ADJ_NPARAM = 9
};

// Per adjoint workgroup partial sums: the parameters, then for each wall of
// the workgroup's simulation the cotangents of its position and velocity
// that the particles' wall contacts produce.
constexpr int ADJ_NPART = ADJ_NPARAM + 2 * MAXWALLS;
// Particles per adjoint workgroup (four work-items each).
constexpr int ADJ_PPB = 64;
// Bond parameters with a cotangent: k_n, k_t, gamma_n, gamma_t.
constexpr int ADJ_NBPARAM = 4;
// Doubles of wall reverse state per wall: cotangents of the position,
// velocity and acceleration the step leaves, and of the force sum it reads.
constexpr int ADJ_WALL = 4;

// What the forward pass recorded for one step.  The state is in the step's
// sorted order; rows and displacements are original-index rows; indices
// are batch indices.
struct AdjStep {
    const DevParams* ptab;     // per-simulation parameters (device)
    const uint32_t* chunk_sim; // simulation of each 256-particle chunk
    uint32_t nsim;
    uint32_t ntot;             // particles including padding
    const double4* x;          // positions at the step start (radius in w)
    const double4* v;          // velocities (fixvel in w)
    const double4* w;          // angular velocities
    const uint32_t* perm;      // sorted index -> original index
    const double4* dt0;        // ntot x NC tangential displacements before the step
    const uint32_t* srt;       // ntot x NC the partner's sorted index, per contact
    const uint32_t* rows;      // ntot x NC contact rows after the interaction pass
    const uint32_t* prev_rows; // the previous step's rows, or null on the first step
    const double4* wall_nx;    // walls at the step start (all simulations' walls)
    const double4* wall_mvfd;
    int first;                 // the run's first step (acc0 = 0)
    int last;                  // the recorded pass's last step (no displacement cotangent)
    // Bonds (nbond = 0 without): each bond's inputs in this step (x_i, x_j,
    // v_i, v_j, w_i, w_j, delta0, omega0), its particles and whether it
    // applied forces; per simulation whether the bond gather of particle
    // nb0 scanned every bond (a broken bond retargeted onto it).
    uint32_t nbond;
    const double4* bond_in;
    const uint4* bond_idx;     // bx, by, ok, 0
    const uint32_t* bond_full;
    const uint32_t* bond_sim;
    const uint32_t* badj_off;  // CSR of the initial bond list (ntot + 1)
    const uint32_t* badj;
};

// Cotangents, updated step by step from the last to the first; the "out"
// arrays (of the step's outputs) are only read, the "in" arrays (of its
// inputs) written, and the two swap between steps.
struct AdjState {
    const double4* Lout[3];  // of the step's outputs x', v', w'
    double4* Lin[3];         // of its inputs x, v, w
    const double4* La_out;   // of the accelerations the step writes (next step's acc0)
    const double4* Laa_out;
    double4* La_in;          // of the accelerations it read (acc0)
    double4* Laa_in;
    const double4* Ldt_in;   // ntot x NC: of the displacements the step writes
    double4* Ldt_out;        // ntot x NC: of the displacements the step read, at the
                             //   slots occupied in prev_rows
    double* part;            // per adjoint workgroup: ADJ_NPART partial sums
    double* gparam;          // nsim x ADJ_NPARAM, accumulated over the steps
    double* wall;            // total walls x ADJ_WALL reverse state
    uint32_t* asym;          // contacts missing from their partner's row
    // bonds
    const double4* Lb_out;   // nbond x 2: of the bond displacements / rotations written
    double4* Lb_in;          // nbond x 2: of those read
    double4* bcontrib;       // nbond x 6: cotangents of x, v, w of particle i then j
    double* bpart;           // nbond x ADJ_NBPARAM
};

// (Σnp x 3) host rows <-> the batch's double4 arrays (w = 0): simulation s's
// particle k is host row ptab[s].hbase + k and batch index ptab[s].base + k.
hipError_t launch_adj_expand(double4* dst, const double* src, const AdjStep& st, hipStream_t s);
hipError_t launch_adj_pack(double* dst, const double4* src, const AdjStep& st, hipStream_t s);
// dst[i].xyz += the host row of particle i.
hipError_t launch_adj_accumulate(double4* dst, const double* src, const AdjStep& st,
                                 hipStream_t s);
// Reverse of one step: particles (contacts, walls, integration), then the
// bonds and their per-particle gather, then per simulation the fixed-order
// sums of the partials and the reverse of the walls' equation of motion.
hipError_t launch_adj_step(const AdjStep& st, const AdjState& g, hipStream_t s);

}  // namespace sphx

#endif
