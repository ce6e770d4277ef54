// Types shared by the host engine and the HIP kernels.
//
// The parameter block replaces the reference's __constant__ singletons
// (devC_params, devC_grid, devC_np, devC_nw, devC_dt; src/constants.cuh:10-17,
// src/datatypes.h:45-101) with one struct passed by value to every kernel.
#ifndef SPHERE_AMD_DEM_TYPES_HPP_
#define SPHERE_AMD_DEM_TYPES_HPP_

#include <stdint.h>

namespace sphx {

constexpr int NC = 32;          // contacts per particle, src/constants.h:21
constexpr int MAXWALLS = 6;     // src/constants.h:12
constexpr int BLOCK = 256;      // particles per workgroup in the per-particle kernels

// A contact-row word packs the partner's index with the periodic image it
// was found through.  On the device the index is the partner's position in
// the order the rows were last written in (the sorted order of the step
// that wrote them, `perm_rows`): the contact kernel reads its partner's
// state there without a lookup, and a step that follows a re-sort maps
// each word to the new order through the original index.  The accessors
// translate to the reference's original indices.  The image shift
// components are exactly 0.0, +L or -L (src/contactsearch.cuh:17-58
// adds/subtracts L to a zero vector), so two bits per periodic axis
// reproduce them bit for bit.
constexpr uint32_t CW_IDX_BITS = 28;
constexpr uint32_t CW_IDX_MASK = (1u << CW_IDX_BITS) - 1u;
constexpr uint32_t CW_SHIFT_X = 28;   // 0: none, 1: +L[0], 2: -L[0]
constexpr uint32_t CW_SHIFT_Y = 30;   // 0: none, 1: +L[1], 2: -L[1]
constexpr uint32_t MAX_NP = 1u << 27; // the reference's own u32 row index limit

// Candidate (Verlet) lists: per particle, in sorted order, the neighbours
// within r_a + r_b + skin found by the 27-cell scan, kept in scan order.
// An entry is {sorted index | cell offset << 27, packed contact word}.
// Per particle the list keeps {count | flags, hit mask}: the hit mask has
// bit RMAX q + i set when entry q + LPP i (the entry lane q of a group of
// LPP lanes tests in its round i, RMAX = NV / LPP) held a contact after the
// last step, and VC_HP_VALID says that the row then held exactly those
// partners.
constexpr int NV = 32;                   // entries per particle
constexpr uint32_t VL_OVERFLOW = 0xffffffffu;
constexpr uint32_t VL_IB_MASK = (1u << 27) - 1u;
constexpr uint32_t VL_CELL_SHIFT = 27;
constexpr uint32_t VC_COUNT_MASK = 0xffffu;
constexpr uint32_t VC_HP_VALID = 1u << 16;

// Particles that changed cell, listed by the integration pass for the
// one-workgroup re-sort at the end of the step.
constexpr uint32_t MV_CAP = 2048;

// An empty contact slot (the reference stores np there; np differs between
// the simulations of a batch, so the engine stores this marker and the
// field accessors translate).
constexpr uint32_t EMPTY = CW_IDX_MASK;

struct DevParams {
    // grid (src/datatypes.h:45-51)
    double origo[3];
    double L[3];
    uint32_t num[3];
    int32_t periodic;
    int32_t adaptive;
    uint32_t ncells;
    // sizes
    uint32_t np;
    uint32_t nw;
    uint32_t nb0;
    uint32_t nblocks;     // ceil(np / 256): the reference's blocksPerGrid
    double dt;
    // physical parameters (src/datatypes.h:74-101)
    double g[3];
    double k_n, k_t, k_r, E;
    double gamma_n, gamma_t, gamma_r;
    double mu_s, mu_d, mu_r;
    double gamma_wn, gamma_wt, mu_ws, mu_wd;
    double rho;
    uint32_t contactmodel;
    double kappa, db, V_b;
    double lambda_bar, sigma_b, tau_b;
    double sigma0_A, sigma0_f;
    double tau_x;
    int32_t wmode[MAXWALLS];
    // candidate lists
    double vl_skin;            // list radius margin (m); 0 disables the lists
    double vl_trigger2;        // squared displacement that forces a rebuild
    int32_t vl_enabled;
    uint32_t vl_cap;           // list capacity in use (<= NV); beyond it the full scan
    // Largest number of particles changing cell that the one-workgroup
    // re-sort merges into the previous order (<= MV_CAP); beyond it, it
    // sorts from scratch.
    uint32_t mv_max;
    // Place of this simulation in the batch arrays.  Its particles occupy
    // original and sorted indices [base, base + npad): the np real ones,
    // then padding particles whose key pad_key sorts after every real key
    // of the simulation and before the next simulation's keys.
    uint32_t sim;
    uint32_t base;
    uint32_t npad;
    uint32_t chunk_base;  // base / 256: the first block partial of the simulation
    uint32_t cell_base;   // global key of the simulation's cell 0
    uint32_t pad_key;
    uint32_t wall_base;   // first wall of the simulation
    uint32_t bond_base;   // first bond of the simulation
    uint32_t rhist_base;  // radix histogram space of the one-workgroup re-sort
    uint32_t wall_sums;   // 1: wall w integrates its own force sum; 0: wall 0's (P8)
    uint32_t hbase;       // first row of the simulation in host arrays that
                          // concatenate the simulations' particles
    // Bond options (0 = the reference's bondsLinear): a bond stiffness in
    // N/m in place of k_n A and k_t A (the bending and twisting stiffness
    // per unit area becoming bond_k / A), and a break at a normal
    // extension of bond_break_ext times the smaller radius in place of the
    // stress criterion.
    double bond_k;
    double bond_break_ext;
};

// Loop state kept on the device so that captured step sequences can be
// replayed: time.current, iter, and the rate-state velocity switch
// (src/device.cu:919,993-999,2454-2462,2536-2538).  The per-step flags are
// double-buffered by iteration parity: step k consumes slot k & 1 while its
// integration pass already hashes the new positions into slot (k + 1) & 1.
struct DevLoop {
    double t;
    int64_t iter;
    double grid_Lz;             // L[2] used by the current step (adaptive grid)
    int32_t velocity_state;
    uint32_t error_bits;        // sticky device-side flags (DevFlag)
    uint32_t changed[2];        // some particle's cell key changed: re-sort
    uint32_t rebuild[2];        // candidate lists must be rebuilt
    uint32_t force_rebuild;     // host request: rebuild the lists this step
    uint32_t zero_empty;        // zero delta_t of empty slots this step
    uint32_t bond_retargeted;   // some broken bond now names particle nb0
    uint32_t vl_builds;         // number of list rebuilds so far (diagnostic)
    uint32_t sorts;             // number of re-sorts so far (diagnostic)
    uint32_t nmovers;           // particles whose key changed for the next step
    uint32_t resorted_last;     // the last step ended with the next step's re-sort
    uint32_t order_changed;     // the order changed since the contact rows were written
    uint32_t copies_stale;      // the sorted copies are in the previous order (read
                                // the original-index arrays through perm instead)
    uint32_t pad1;
};

// Batch-wide state.
struct DevGlobal {
    uint32_t changed[2];  // some simulation needs a re-sort (per step parity)
};

enum DevFlag : uint32_t {
    DFLAG_KEY_RANGE = 1u,          // a cell key exceeds the sorted bit range
    DFLAG_CONTACT_OVERFLOW = 2u,   // more than NC overlaps for a particle (P5);
                                   // informational, the reference's behaviour
    DFLAG_NONFINITE = 4u,          // NaN position at hashing (src/sorting.cuh:39-43)
};

}  // namespace sphx

#endif
