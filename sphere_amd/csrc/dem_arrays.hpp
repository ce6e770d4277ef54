// Device array set of one simulation context.  All particle arrays are
// indexed by original particle index (the .bin order) unless named *_s
// (sorted order), mirroring the reference's dev_* / dev_*_sorted split
// (src/datatypes.h:16-61).
#ifndef SPHERE_AMD_DEM_ARRAYS_HPP_
#define SPHERE_AMD_DEM_ARRAYS_HPP_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dem_types.hpp"

namespace sphx {

// All simulations of a context share these arrays: simulation s owns the
// particle range [ptab[s].base, ptab[s].base + ptab[s].npad), which is a
// whole number of 256-particle chunks, its walls, bonds and cells.
// Parts of a particle's output record: force (w zero), torque (w zero),
// es_dot, ev_dot, p and its force on wall 0, its forces on walls 1-4.
constexpr uint32_t OUT_F = 0, OUT_T = 1, OUT_E = 2, OUT_W = 3;
// The walls with a force per particle (top, two x walls, two y walls).
constexpr uint32_t OUT_NWALL = 5;

struct DevArrays {
    // batch layout
    const DevParams* ptab;      // per-simulation parameters
    const uint32_t* chunk_sim;  // simulation of each 256-particle chunk
    const uint32_t* bond_sim;   // simulation of each bond
    DevGlobal* glob;
    uint32_t nsim;
    uint32_t ntot;              // particles including padding (= 256 * nchunk)
    uint32_t nchunk;
    uint32_t nbond;
    uint32_t nkeys;             // global cell keys (length of cell_start/cell_end)
    uint32_t xcd_order;         // 1: contiguous workgroup ranges per XCD (kernels that gather),
                                // 0: hardware order
    // kinematics, 32 B per particle each (x,y,z,w)
    double4* x;        // position, radius in w
    double4* vel;      // velocity, fixvel in w
    double4* angvel;
    double4* angpos;
    double4* xyzsum;
    double4* acc;      // previous-step acceleration (TY3 acc0)
    double4* angacc;
    // The contact pass's outputs: one 128-byte record (four double4 parts,
    // OUT_*) per particle, written whole by the lanes of a contact group
    // (scattered by original index, each record is one cache line).  The
    // energies es and ev they feed (es += es_dot dt) are advanced by the
    // integration pass, in original order.
    double4* out;
    double* es;
    double* ev;
    // sorting
    uint32_t* key;       // cell key per original index (hash output)
    uint32_t* key_prev;  // key of the last sorted step, per original index
    uint32_t* skey;      // sorted keys
    uint32_t* perm;      // sorted index -> original index
    uint32_t* key_tmp;   // radix ping-pong
    uint32_t* val_tmp;
    uint32_t* rhist;     // radix digit histograms, 256 x tiles
    uint32_t* cell_start;
    uint32_t* cell_end;
    // Sorted copies of x, vel and angvel: the contact kernel reads its own
    // and its partners' state in sorted order, which keeps those gathers
    // local.  Two sets, by loop-counter parity: step k reads set k & 1 and
    // whatever writes the state for step k + 1 (the integration, a re-sort
    // at the end of step k) writes set (k + 1) & 1, so that a kernel may
    // integrate particles while others still read their step-k state.
    double4* xs[2];
    double4* vs[2];
    double4* ws[2];
    uint32_t* rank;      // original index -> sorted index (inverse of perm)
    uint32_t* movers;    // per simulation, MV_CAP indices whose key changed (unordered)
    // The order a step used, kept when the step's end already re-sorted for
    // the next one (what the reference's arrays hold after that step).
    uint32_t* skey_bk;
    uint32_t* perm_bk;
    uint32_t* cs_bk;
    uint32_t* ce_bk;
    // candidate lists (sorted-index rows of NV entries)
    uint2* vlist;
    uint2* vcount;       // {entries per row | flags, or VL_OVERFLOW; hit mask}
    double4* x_build;    // position at the last list build, original index
    // contact history (NC words per particle, original-index rows)
    uint32_t* crow;      // packed partner index (in perm_rows' order) + image code
    uint32_t* perm_rows; // the order the row words index: sorted index -> original
    double4* delta_t;    // tangential displacement per slot (w unused)
    // Recording for the adjoint (null otherwise): the contact pass writes
    // its rows and, per contact, the displacement before the step and the
    // partner's sorted index (original-index rows); the state the step
    // starts from and the order, in sorted order; the walls.  The bond pass
    // writes each bond's inputs (x, v, w of both particles, displacement
    // and rotation; 8 per bond), its particles and whether it applied
    // forces, and the bond gather whether particle nb0 scanned every bond.
    uint32_t* tape_rows;
    double4* tape_dt0;
    uint32_t* tape_srt;
    uint32_t* tape_perm;
    double4* tape_x;
    double4* tape_v;
    double4* tape_w;
    double4* tape_wnx;
    double4* tape_wmv;
    double4* tape_bond_in;
    uint4* tape_bond_idx;
    uint32_t* tape_bond_full;
    // walls
    double4* w_nx;       // normal (xyz) and position (w)
    double4* w_mvfd;     // mass, velocity, force, sigma0
    double* w_acc;
    double* tau_pp;      // np shear stress per particle (wmode 3)
    double* partial;     // block sums, one per 256-particle chunk
    double* tau_partial;
    double* tau_eff;     // serial sum of the tau partials, one per simulation
    // bonds
    uint32_t* bonds;     // nb0 x 2
    double4* bonds_delta;
    double4* bonds_omega;
    double4* bond_f;     // per-bond force on particle i (this step)
    double4* bond_t;     // per-bond torque t_j (this step)
    uint32_t* bond_ok;   // 1 if the bond applied forces this step
    uint32_t* badj_off;  // CSR offsets (np + 1) of bonds per particle
    uint32_t* badj;      // CSR bond ids, ascending per particle
    // loop state, one per simulation
    DevLoop* loop;
};

}  // namespace sphx

#endif
