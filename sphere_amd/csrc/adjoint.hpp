// Reverse-mode step (contact model 2) over a recorded forward pass.
#ifndef SPHERE_AMD_ADJOINT_HPP_
#define SPHERE_AMD_ADJOINT_HPP_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dem_types.hpp"

namespace sphx {

// Parameters with a cotangent, in output order.
enum AdjParam : int {
    ADJ_KN = 0,    // k_n (contacts and walls)
    ADJ_KT = 1,    // k_t
    ADJ_GN = 2,    // gamma_n
    ADJ_GT = 3,    // gamma_t
    ADJ_MUS = 4,   // mu_s
    ADJ_MUD = 5,   // mu_d
    ADJ_GWN = 6,   // gamma_wn
    ADJ_GWT = 7,   // gamma_wt
    ADJ_MUWD = 8,  // mu_wd
    ADJ_NPARAM = 9
};

// What the forward pass recorded for one step.  The state is in the step's
// sorted order; rows and displacements are original-index rows.
struct AdjStep {
    const double4* x;          // positions at the step start (radius in w)
    const double4* v;          // velocities (fixvel in w)
    const double4* w;          // angular velocities
    const uint32_t* perm;      // sorted index -> original index
    const double4* dt0;        // np x NC tangential displacements before the step
    const uint32_t* srt;       // np x NC the partner's sorted index, per contact
    const uint32_t* rows;      // np x NC contact rows after the interaction pass
    const uint32_t* prev_rows; // the previous step's rows, or null on the first step
    const double4* wall_nx;    // walls at the step start
    const double4* wall_mvfd;
    int first;                 // the run's first step (acc0 = 0)
    int last;                  // the recorded pass's last step (no displacement cotangent)
};

// Cotangents, updated step by step from the last to the first.
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
    const double4* Ldt_in;   // np x NC: of the displacements the step writes
    double4* Ldt_out;        // np x NC: of the displacements the step read, at the
                             //   slots occupied in prev_rows
    double* gparam;          // ADJ_NPARAM sums
    uint32_t* asym;          // contacts missing from their partner's row
};

// (n x 3) doubles <-> n double4 (w = 0).
hipError_t launch_adj_expand(double4* dst, const double* src, uint32_t n, hipStream_t s);
hipError_t launch_adj_pack(double* dst, const double4* src, uint32_t n, hipStream_t s);
// dst[i].xyz += the i-th row of the (n x 3) src.
hipError_t launch_adj_accumulate(double4* dst, const double* src, uint32_t n, hipStream_t s);
hipError_t launch_adj_step(const DevParams& p, const AdjStep& st, const AdjState& g, double Lz,
                           hipStream_t s);

}  // namespace sphx

#endif
