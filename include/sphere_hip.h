/*
 * sphere_hip: MI355X-native DEM particle timestep for sphere, C ABI.
 *
 * The reference has no in-process API: its boundary is a process
 * (python/sphere.py:4482-4540 writes input/<sid>.bin and runs
 * `./sphere [-q] [--dry] [-d N] input/<sid>.bin`, src/main.cpp:28-223) plus the
 * versioned .bin format (src/file_io.cpp:30-731).  Each entry point below
 * replaces one piece of that boundary; the citation names the reference
 * code whose behaviour it carries.
 *
 * Conventions: plain pointers and sizes only; 0 = success, negative = error
 * (the library never calls exit()); sph_last_error() returns the message.
 * A context owns all device memory and one HIP stream; one host thread per
 * context.
 */
#ifndef SPHERE_HIP_H_
#define SPHERE_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Maximum contacts per particle (src/constants.h:21) and walls (:12). */
#define SPH_NC 32
#define SPH_MAXWALLS 6
/* The .bin version this library reads and writes (src/version.h:5). */
#define SPH_VERSION 2.15

/* Error codes. */
#define SPH_OK 0
#define SPH_E_IO (-1)        /* file cannot be opened / short read        */
#define SPH_E_VERSION (-2)   /* .bin version != 2.15 (src/file_io.cpp:49-54) */
#define SPH_E_INPUT (-3)     /* checkValues failure (src/sphere.cpp:145-345) */
#define SPH_E_DEVICE (-4)    /* HIP runtime error                           */
#define SPH_E_ARG (-5)       /* bad argument (field id, size, index)        */
#define SPH_E_UNSUPPORTED (-6)

/* Field ids for sph_get_field / sph_set_field.  Particle fields are indexed
 * by original particle index (the .bin order), as the reference's dev_* arrays
 * are.  Layouts (little-endian, row-major):
 *   X         np x 4 f64 (x, y, z, radius)         k.x        src/sphere.h:63
 *   VEL       np x 4 f64 (vx, vy, vz, fixvel)      k.vel
 *   ANGVEL .. ANGACC  np x 3 f64
 *   ES_DOT .. P       np f64
 *   CONTACTS  np x NC u32 (original index of the partner, np = empty slot)
 *   DELTA_T   np x NC x 3 f64 (accumulated tangential displacement)
 *   DISTMOD   np x NC x 4 f64 (periodic shift x,y,z; radius of the partner)
 *   CELLID    np u32 sorted cell keys; PERM np u32 sorted index -> original
 *   CELLSTART/CELLEND  ncells u32 (cellEnd is meaningful only where
 *             cellStart != 0xffffffff, src/device.cu:1089-1090)
 *   WALLS_NX, WALLS_MVFD  nw x 4 f64; WALLS_ACC nw f64; WALLS_FORCE_PP nw x np
 *   BONDS nb0 x 2 u32; BONDS_DELTA, BONDS_OMEGA nb0 x 4 f64 (t.x,t.y,t.z,n)
 *   COLOR np i32
 */
enum sph_field {
    SPH_F_X = 0,
    SPH_F_VEL = 1,
    SPH_F_ANGVEL = 2,
    SPH_F_ANGPOS = 3,
    SPH_F_FORCE = 4,
    SPH_F_TORQUE = 5,
    SPH_F_XYZSUM = 6,
    SPH_F_ACC = 7,
    SPH_F_ANGACC = 8,
    SPH_F_ES_DOT = 9,
    SPH_F_ES = 10,
    SPH_F_EV_DOT = 11,
    SPH_F_EV = 12,
    SPH_F_P = 13,
    SPH_F_CONTACTS = 14,
    SPH_F_DELTA_T = 15,
    SPH_F_DISTMOD = 16,
    SPH_F_CELLID = 17,
    SPH_F_PERM = 18,
    SPH_F_CELLSTART = 19,
    SPH_F_CELLEND = 20,
    SPH_F_WALLS_NX = 21,
    SPH_F_WALLS_MVFD = 22,
    SPH_F_WALLS_ACC = 23,
    SPH_F_WALLS_FORCE_PP = 24,
    SPH_F_BONDS = 25,
    SPH_F_BONDS_DELTA = 26,
    SPH_F_BONDS_OMEGA = 27,
    SPH_F_COLOR = 28,
    SPH_F_COUNT = 29
};

/* Scalar state ids for sph_get_scalar / sph_set_scalar. */
enum sph_scalar {
    SPH_S_TIME = 0,        /* time.current  */
    SPH_S_TIME_TOTAL = 1,  /* time.total    */
    SPH_S_DT = 2,          /* time.dt       */
    SPH_S_FILE_DT = 3,     /* time.file_dt  */
    SPH_S_ITER = 4,        /* loop counter (restarts at 0 per run, src/device.cu:919) */
    SPH_S_STEP_COUNT = 5,  /* time.step_count */
    SPH_S_NP = 6,
    SPH_S_NW = 7,
    SPH_S_NB0 = 8,
    SPH_S_NCELLS = 9,
    SPH_S_VL_SKIN = 10,    /* candidate-list skin (m); 0 = scan 27 cells every step */
    SPH_S_SORTS = 11,      /* re-sorts performed so far (diagnostic, read-only)      */
    SPH_S_VL_BUILDS = 12,  /* candidate-list rebuilds so far (diagnostic, read-only) */
    SPH_S_NSIMS = 13,      /* simulations in the context (read-only) */
    SPH_S_CONTACT_LANES = 14, /* work-items per particle in the contact kernel for
                                 contact models 2 and 3: 1, 4 (default), 8 or 16.
                                 Results are identical for every value. */
    SPH_S_SORT_MODE = 15,   /* where the re-sort runs: 0 = sort launches at each step
                               start, 1 = one workgroup at the end of the previous step
                               when some particle changed cell, 2 (default) = chosen
                               from the measured re-sort rate.  Identical results. */
    SPH_S_RESORT_MERGE_MAX = 16, /* particles changing cell up to which the end-of-step
                                    re-sort merges them into the previous order
                                    (0..2048, default 2048); above it, it sorts anew. */
    SPH_S_GRAPH_STEPS = 17, /* > 0: sph_step replays steady-state chunks of this many
                               steps from a captured HIP graph (default 0: launches). */
    SPH_S_XCD_ORDER = 18,   /* 1: the gathering kernels hand each XCD one contiguous
                               range of workgroups; 0 (default): hardware order
                               (round-robin over the XCDs), measured faster on the
                               sheared bed.  Identical results. */
    SPH_S_ADJ_CHECKPOINT = 19, /* sph_adjoint: record the forward pass in segments of
                                 this many steps from checkpoints of the whole
                                 context, recomputing each segment before its
                                 reverse (memory for K steps plus one checkpoint
                                 per segment); 0 (default): record every step,
                                 or segments if that does not fit in memory.
                                 Identical gradients either way. */
    SPH_S_FUSED = 20,         /* 1: where the step allows (fixed grid height, no
                                 bonds, no wmode-3 wall, 4 contact lanes, model 2
                                 or 3) the contact kernel also integrates each
                                 particle (one vector component per lane) and a
                                 small pass forms the wall-force partials; 0
                                 (default, measured faster): separate integration
                                 kernel.  Identical results. */
    SPH_S_WALL_SUMS = 21,     /* per simulation.  0 (default, the reference,
                                 src/device.cu:2504-2505): every dynamic wall
                                 integrates wall 0's force sum; 1: wall w
                                 integrates the sum of its own per-particle
                                 forces (same tree and serial order) - the
                                 correction multi-wall setups (triaxial, nw = 5)
                                 need.  No counterpart in the reference. */
    SPH_S_K_N = 22,           /* per simulation, read-only: the normal stiffness
                                 params.k_n, which DEM::forcechains scales the
                                 overlaps by (src/sphere.cpp:789). */
    SPH_S_INTEGRATE_LANES = 23,/* work-items per particle in the integration
                                 pass: 1 (default) or 4 (one vector component
                                 each, then a wall-partials pass; used with a
                                 fixed grid height and no wmode-3 wall;
                                 measured slower on the sheared bed, 22 + 6
                                 us against 18).  Identical results. */
    SPH_S_BOND_K = 24,        /* per simulation.  0 (default, the reference,
                                 src/cohesion.cuh:124,177): bonds take k_n A and
                                 k_t A as stiffness; > 0: this bond stiffness
                                 in N/m for both (bending and twisting
                                 stiffness per unit area bond_k / A).  Not in
                                 the .bin (fixed layout): set after opening. */
    SPH_S_BOND_BREAK_EXT = 25,/* per simulation.  0 (default, the reference,
                                 src/cohesion.cuh:193-206): bonds break on the
                                 stress criterion (sigma_b, tau_b); > 0: a bond
                                 breaks when its normal extension reaches this
                                 fraction of the smaller radius.  Not in the
                                 .bin: set after opening. */
    SPH_S_CONTACT_WG = 26,    /* work-items per workgroup of the contact kernel
                                 with 4 lanes per particle: 128, 256 (default)
                                 or 512.  Identical results. */
    SPH_S_VL_CAP = 27,        /* candidate-list capacity per particle, 8..32
                                 (default 32): a particle with more candidates
                                 scans its 27 cells every step.  Identical
                                 results. */
    SPH_S_RADIX_BITS = 28     /* digit width of the radix sort: 4, 6 or 8
                                 (default); the passes cover the keys' bit
                                 length.  Identical results. */
};

/* Open flags. */
#define SPH_OPEN_QUIET 0x1u      /* no stdout progress lines                     */
#define SPH_OPEN_NOCHECK 0x2u    /* skip checkValues (src/sphere.cpp:145-345)     */
#define SPH_OPEN_SIDECAR 0x4u    /* restart sidecars: sph_open reads "<name>.state"
                                    next to "<name>.bin" when present, and sph_run
                                    writes one next to every output .bin (no
                                    counterpart in the reference, whose restarts
                                    lose the contact history and acc0) */

typedef struct sph_ctx sph_ctx;

/* readbin + checkValues + H2D: DEM::DEM(inputbin, ...) with initCuda=1,
 * src/sphere.cpp:16-95 and src/file_io.cpp:30-418.  device < 0 = current. */
int sph_open(const char* bin_path, int device, unsigned flags, sph_ctx** out);
/* Same, from a .bin image already in host memory. */
int sph_open_mem(const void* bytes, size_t nbytes, int device, unsigned flags,
                 sph_ctx** out);

/* Batched contexts (no counterpart in the reference, whose binary runs one
 * simulation per process): nsims independent simulations, each from its own
 * .bin (own np, grid, walls, bonds, parameters and dt; one contact model for
 * all), stepped together by the same kernel launches.  Simulation k's
 * results are identical to opening its .bin alone.  The plain entry points
 * act on simulation 0; the *_sim variants take the simulation index.  ITER
 * and the performance scalars apply to the whole context. */
int sph_open_batch(const char* const* bin_paths, int nsims, int device, unsigned flags,
                   sph_ctx** out);
int sph_open_batch_mem(const void* const* images, const size_t* sizes, int nsims, int device,
                       unsigned flags, sph_ctx** out);
int sph_num_sims(const sph_ctx* ctx);

/* Run nsteps iterations of the fluid-free loop body of DEM::startTime,
 * src/device.cu:1031-1200,2423-2543, without file output and without the
 * t <= t_total test.  Asynchronous w.r.t. the host until a get/sync call. */
int sph_step(sph_ctx* ctx, uint64_t nsteps);
/* Block until all queued steps are complete. */
int sph_sync(sph_ctx* ctx);

/* DEM::startTime (src/device.cu:826-2700): writes out_dir/<sid>.output00000.bin
 * when step_count == 0, steps while t <= t_total, writes
 * out_dir/<sid>.output%05d.bin every file_dt and out_dir/<sid>.status.dat. */
int sph_run(sph_ctx* ctx, const char* out_dir, const char* sid);

/* DEM::writebin (src/file_io.cpp:420-731) of the current state. */
int sph_write_bin(sph_ctx* ctx, const char* path);
int sph_write_bin_sim(sph_ctx* ctx, int sim, const char* path);
/* Same, to memory.  Returns the image size in *nbytes; writes only when
 * cap >= that size (call with dst = NULL to query). */
int sph_write_mem(sph_ctx* ctx, void* dst, size_t cap, size_t* nbytes);
int sph_write_mem_sim(sph_ctx* ctx, int sim, void* dst, size_t cap, size_t* nbytes);

/* D2H / H2D of one field (layouts above); bytes must equal the field size. */
int sph_get_field(sph_ctx* ctx, int field, void* host_dst, size_t bytes);
int sph_set_field(sph_ctx* ctx, int field, const void* host_src, size_t bytes);
int64_t sph_field_bytes(sph_ctx* ctx, int field);
int sph_get_scalar(sph_ctx* ctx, int which, double* value);
int sph_set_scalar(sph_ctx* ctx, int which, double value);
int sph_get_field_sim(sph_ctx* ctx, int sim, int field, void* host_dst, size_t bytes);
int sph_set_field_sim(sph_ctx* ctx, int sim, int field, const void* host_src, size_t bytes);
int64_t sph_field_bytes_sim(sph_ctx* ctx, int sim, int field);
int sph_get_scalar_sim(sph_ctx* ctx, int sim, int which, double* value);
int sph_set_scalar_sim(sph_ctx* ctx, int sim, int which, double value);

/* Bit-exact check hooks (original-index keyed, as the reference's
 * dev_contacts / dev_delta_t / dev_distmod, src/contactsearch.cuh:273-291). */
int sph_get_contacts(sph_ctx* ctx, uint32_t* idx, double* delta_t, double* distmod);
int sph_get_cells(sph_ctx* ctx, uint32_t* cell_id, uint32_t* perm,
                  uint32_t* cell_start, uint32_t* cell_end);

/* Adjoint (reverse mode) of nsteps steps; no counterpart in the reference.
 * Runs the steps forward from the current state, recording each step's
 * inputs, contact rows and bonds, then backwards with that contact list and
 * every branch (force clamps, stick or slip, wall friction, bond breakage)
 * frozen, and returns the gradient of
 *     L = sum_i dL_dx[i] . x_i + dL_dv[i] . v_i + dL_dw[i] . w_i
 * (state after the steps) with respect to the state before them (g_x0,
 * g_v0, g_w0: np x 3 each) and to the parameters (g_params, 9 values:
 * k_n, k_t, gamma_n, gamma_t, mu_s, mu_d, gamma_wn, gamma_wt, mu_wd).  Any
 * pointer may be NULL.  Lengths are differentiated as exact roots.  The
 * reverse covers contactLinear (src/contactmodels.cuh:244-411),
 * contactLinear_wall (:10-117), bondsLinear (src/cohesion.cuh:8-233), the
 * particles' TY3 integration (src/integration.cuh:14-305) and the walls'
 * equation of motion for fixed, stress-controlled and velocity-controlled
 * walls (wmode 0, 1, 2; src/integration.cuh:344-450).  Not covered
 * (SPH_E_UNSUPPORTED): contact models 1 and 3, cohesion, E-scaled
 * stiffness, wmode 3, adaptive grids, rate-state particles.  Parameter
 * sums are formed in a fixed order: gradients are reproducible bit for
 * bit.  Single-simulation contexts (see sph_adjoint_batch).  The context
 * is left at the state after the steps. */
int sph_adjoint(sph_ctx* ctx, uint64_t nsteps, const double* dL_dx, const double* dL_dv,
                const double* dL_dw, double* g_x0, double* g_v0, double* g_w0,
                double* g_params);

/* Adjoint of a trajectory loss; no counterpart in the reference.  Runs
 * every * nsamples steps forward from the current state like sph_adjoint
 * and returns the gradient of
 *     L = sum_k sum_i dL_dx[k][i] . x_i(k) + dL_dv[k][i] . v_i(k)
 * where x(k), v(k) is the state after (k + 1) * every steps; dL_dx and dL_dv
 * are nsamples x np x 3 (either may be NULL).  Each term enters the reverse
 * pass at its step, so the cotangents of the accelerations and of the
 * contact displacements carry across the sampling times.  Same coverage,
 * outputs and final state as sph_adjoint. */
int sph_adjoint_samples(sph_ctx* ctx, uint64_t every, uint64_t nsamples, const double* dL_dx,
                        const double* dL_dv, double* g_x0, double* g_v0, double* g_w0,
                        double* g_params);

/* The same two adjoints over every simulation of a context (a batch of
 * independent simulations, e.g. one calibration per simulation); no
 * counterpart in the reference.  Particle arrays (dL_dx, dL_dv, dL_dw and
 * g_x0, g_v0, g_w0; for the samples variant nsamples of them) concatenate
 * the simulations' particles in simulation order (sum of np rows x 3);
 * g_params receives nsims x 9 values, simulation by simulation, each the
 * gradient with respect to that simulation's own parameters.  On a
 * single-simulation context they equal sph_adjoint / sph_adjoint_samples. */
int sph_adjoint_batch(sph_ctx* ctx, uint64_t nsteps, const double* dL_dx, const double* dL_dv,
                      const double* dL_dw, double* g_x0, double* g_v0, double* g_w0,
                      double* g_params);
int sph_adjoint_samples_batch(sph_ctx* ctx, uint64_t every, uint64_t nsamples,
                              const double* dL_dx, const double* dL_dv, double* g_x0,
                              double* g_v0, double* g_w0, double* g_params);

/* sph_adjoint_samples (and its batched form) in two calls, for callers
 * that need the forward pass's samples before the loss is known (a
 * torch.autograd forward, then its backward); no counterpart in the
 * reference.  sph_adjoint_record runs nsamples x every steps from the
 * current state recording the tape (or checkpoints, as sph_adjoint does),
 * and writes the positions after every, 2 every, ... steps to samples_x
 * (nsamples x rows x 3, the simulations' particles concatenated; may be
 * NULL).  sph_adjoint_reverse then takes the loss gradients of those
 * samples (dL_dx, dL_dv as in sph_adjoint_samples) and gives the same
 * gradients, bit for bit, without stepping forward again (with checkpoints
 * it recomputes each segment).  The context must neither step nor record
 * again in between (SPH_E_ARG); the state after the reverse is the one
 * after the forward pass. */
int sph_adjoint_record(sph_ctx* ctx, uint64_t every, uint64_t nsamples, double* samples_x);
int sph_adjoint_reverse(sph_ctx* ctx, uint64_t every, uint64_t nsamples, const double* dL_dx,
                        const double* dL_dv, double* g_x0, double* g_v0, double* g_w0,
                        double* g_params);

/* Restart sidecar (no counterpart in the reference: src/file_io.cpp:420-731
 * writes none of this, so a run resumed from an output .bin starts with
 * empty contact rows, zero displacements and acc0 = 0).  Holds the contact
 * rows with their periodic images, the tangential displacements, the
 * accelerations of the last step, the walls' accelerations and the loop
 * counter, so that a run resumed from an output .bin and its sidecar is
 * bit-identical to the uninterrupted one.  Reading checks np, nw and NC. */
int sph_write_state(sph_ctx* ctx, const char* path);
int sph_read_state(sph_ctx* ctx, const char* path);
int sph_write_state_sim(sph_ctx* ctx, int sim, const char* path);
int sph_read_state_sim(sph_ctx* ctx, int sim, const char* path);

/* Per-kernel timing: mean device microseconds of the last sph_step call's
 * kernels, measured with HIP events on the context stream
 * (the reference's PROFILING timers, src/device.cu:947-987,2711-2808). */
/* Online autotuner; no counterpart in the reference.  A coordinate search,
 * one knob at a time from the incumbent settings, keeping a value that is
 * at least 3% faster: contact lanes (4, 8, 16; contact models 2 and 3), the
 * contact workgroup size (128, 256, 512; 4 lanes), the candidate-list
 * capacity (16, 24, 32) and the workgroup order (SPH_S_XCD_ORDER), each
 * judged by the contact kernel's mean time over steps_per_trial steps
 * (HIP events) after steps_per_trial / 2 untimed steps; then the radix
 * digit width (4, 6, 8 bits), judged by four full sorts of the current keys.
 * Every setting gives the same results bit for bit, so the trials are
 * ordinary steps: the simulation advances by 1.5 x steps_per_trial steps
 * per contact trial (up to 11).  *us_per_step (may be NULL) receives the
 * contact kernel's best mean time in microseconds; SPH_S_CONTACT_LANES,
 * SPH_S_CONTACT_WG, SPH_S_VL_CAP, SPH_S_RADIX_BITS and SPH_S_XCD_ORDER then
 * read the chosen settings. */
int sph_autotune(sph_ctx* ctx, uint64_t steps_per_trial, double* us_per_step);

int sph_profile(sph_ctx* ctx, int enable);
int sph_profile_report(sph_ctx* ctx, char* buf, size_t cap);

void sph_close(sph_ctx* ctx);
const char* sph_last_error(const sph_ctx* ctx);
/* Error message of the last failed sph_open / sph_open_mem on this thread. */
const char* sph_open_error(void);
const char* sph_version_string(void);

/* Output-side readers of the current state, on the GPU.
 * DEM::findOverlaps (src/sphere.cpp:674-724, printed by `sphere --contacts`,
 * src/sphere.cpp:904-913): every pair i < j (.bin indices, ascending i then
 * j) whose FP64 distance is below r_i + r_j, no periodic images, both with
 * fixvel 0; pairs[2k], pairs[2k + 1] and overlaps[k] = distance - (r_i + r_j)
 * for the first min(total, cap) pairs; *count = total.
 * DEM::porosity (src/sphere.cpp:486-592, `sphere --porosity`): z_slices
 * slices from origo to the top wall (or L_z); SPH_E_INPUT when the slices
 * are thinner than twice the last particle's radius (the reference exits). */
int sph_find_overlaps(sph_ctx* ctx, int sim, uint32_t* pairs, double* overlaps, uint64_t cap,
                      uint64_t* count);
int sph_porosity(sph_ctx* ctx, int sim, int z_slices, double* z_pos, double* porosity);

/* No counterpart in the reference.  sph_porosity for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the per-slice results are staged
 * in a fixed device array of 2 x SPH_POROSITY_NOALLOC_MAX_SLICES doubles
 * (1 MiB) that the library holds on each GPU it runs on until the process
 * ends, not in a buffer allocated per call.  Same results and errors as
 * sph_porosity; SPH_E_ARG when z_slices exceeds the maximum.  Every call on
 * one GPU uses that GPU's array: calls that overlap on one GPU (from
 * several host threads, or contexts on different streams) race. */
#define SPH_POROSITY_NOALLOC_MAX_SLICES 65536
int sph_porosity_noalloc(sph_ctx* ctx, int sim, int z_slices, double* z_pos,
                         double* porosity);

/* Diagnostic, no reference counterpart: runs the engine's FP32 square root
 * (the reference's sqrtf in length(), src/vector_arithmetic.h) on device
 * `device` for every FP32 bit pattern in [lo, hi) (hi <= 2^32) and counts
 * the results that differ from the correctly rounded root. */
int sph_selftest_sqrt32(int device, uint64_t lo, uint64_t hi, uint64_t* mismatches);

#ifdef __cplusplus
}
#endif

#endif
