/*
 * CPU oracle for the sphere DEM particle timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * it and never falls back to it.
 *
 * It is a plain C++ restatement of the reference's per-step algorithm
 * (reference files cited per function in sphere_oracle.cpp), with host IEEE
 * semantics: no FMA contraction (-ffp-contract=off), correctly rounded
 * sqrtf for the FP32 length() of src/vector_arithmetic.h:317-324.
 *
 * Input is a sphere .bin v2.15 image (src/file_io.cpp:30-418); fields are
 * read back with orc_get_field() using the SPH_F_* ids of
 * include/sphere_hip.h, laid out exactly as sph_get_field() lays them out.
 */
#ifndef SPHERE_ORACLE_H_
#define SPHERE_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* Parse a v2.15 .bin image; returns NULL and fills err (if given) on error. */
orc_sim* orc_open_mem(const void* bytes, size_t nbytes, char* err, size_t errlen);
void orc_close(orc_sim* s);

/* Run n iterations of the loop body of DEM::startTime (src/device.cu:1031-2543),
 * without file output and without the t <= t_total test. */
int orc_step(orc_sim* s, int64_t nsteps);

/* Run one named stage of the step, for stage-level parity tests. */
enum orc_stage_id {
    ORC_HASH = 0,       /* calcParticleCellID               src/sorting.cuh:25-51 */
    ORC_SORT = 1,       /* thrust::sort_by_key              src/device.cu:1075-1078 */
    ORC_REORDER = 2,    /* cellStart reset + reorderArrays  src/device.cu:1089; src/sorting.cuh:56-132 */
    ORC_TOPOLOGY = 3,   /* topology                         src/contactsearch.cuh:329-371 */
    ORC_INTERACT = 4,   /* interact                         src/contactsearch.cuh:385-726 */
    ORC_BONDS = 5,      /* bondsLinear (serial)             src/cohesion.cuh:8-233 */
    ORC_INTEGRATE = 6,  /* shear-stress BC + integrate      src/integration.cuh:14-305,457-484 */
    ORC_WALLS = 7,      /* summation + integrateWalls       src/integration.cuh:309-450 */
    ORC_ADVANCE = 8     /* t += dt; ++iter                  src/device.cu:2536-2538 */
};
int orc_stage(orc_sim* s, int stage);

/* Copy a field to dst (bytes must equal the field size). Returns 0 on success. */
int orc_get_field(const orc_sim* s, int field, void* dst, size_t bytes);
int64_t orc_field_bytes(const orc_sim* s, int field);
/* Overwrite a field from src (same layout as orc_get_field). */
int orc_set_field(orc_sim* s, int field, const void* src, size_t bytes);

/* Serialize the current state as a v2.15 .bin image (src/file_io.cpp:420-731).
 * Returns the number of bytes needed; writes only if cap is large enough. */
int64_t orc_write_mem(const orc_sim* s, void* dst, size_t cap);

uint32_t orc_np(const orc_sim* s);
uint32_t orc_nc(void);
double orc_time(const orc_sim* s);
int64_t orc_iter(const orc_sim* s);
/* Test hooks: the FP32 length rule on one value (sqrtf((float)v) widened),
 * and the reference's wall-force reduction (256-wide block trees, then a
 * serial sum over blocks) of n values. */
double orc_sqrt32(double v);
/* Test aid for gradient checks: FP64 lengths instead of FP32 (default 0). */
void orc_set_exact_length(int on);
/* Test aid: record, in every following interact stage (contact models 2
 * and 3), each evaluated contact's overlap, normal, forces, tangential
 * velocity and displacements (np x NC records of orc_contact_log_width()
 * doubles, original-index rows; see CLOG_W in sphere_oracle.cpp). */
void orc_set_contact_log(int on);
int orc_contact_log_width(void);
int orc_get_contact_log(const orc_sim* s, double* dst, size_t n);
/* The engine's bond options (SPH_S_BOND_K in N/m, SPH_S_BOND_BREAK_EXT as
 * a fraction of the smaller radius; 0 = the reference's law) for this
 * simulation. */
void orc_set_bond_model(orc_sim* s, double bond_k, double bond_break_ext);
/* Each dynamic wall integrates its own force sum instead of wall 0's (the
 * engine's optional correction of the reference's P8 behaviour). */
void orc_set_wall_sums_per_wall(int on);
double orc_tree_sum(const double* v, uint32_t n);
/* Output-side readers: DEM::findOverlaps (pairs i < j, FP64 lengths, no
 * periodic images, fixvel 0) and DEM::porosity (z slices, -1 when too thin). */
void orc_find_overlaps(const orc_sim* s, uint32_t* pairs, double* delta, uint64_t cap,
                       uint64_t* count);
int orc_porosity(const orc_sim* s, int z_slices, double* z_pos, double* porosity);

/* Number of OpenMP threads the per-particle stages use (0 = runtime default). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif
