/*
 * hugks.h -- C-ABI of the MI355X Hermite-UGKS time-step library
 * (libhugks.so, HIP/gfx950).  Plain C types only; callable from Fortran
 * through ISO_C_BINDING (hermite-ugks_amd/driver/hugks_capi.f90), from C,
 * and from Python ctypes.
 *
 * The reference solver has no FFI: its hot path is four argument-less
 * Fortran subroutines acting on module globals, called once per step by
 * the main program:
 *     timestep()       2d/src/solver.f90:13-61    1d/src/solver.f90:16-47
 *     interpolation()  2d/src/solver.f90:64-122   1d/src/solver.f90:49-65
 *     evolution()      2d/src/solver.f90:161-243  1d/src/solver.f90:106-139
 *     update()         2d/src/solver.f90:245-395  1d/src/solver.f90:141-264
 * called at 2d/src/main.f90:78-82 and 1d/src/main.f90:86-89.  The entry
 * points below replace exactly that seam.  Everything else (namelist,
 * case selection, velocity quadrature, initial condition, output) stays in
 * the caller, which hands over the data those routines read from the
 * reference modules CONFIG/ALLDATA/VELOCITY.
 *
 * Conventions
 *   - Every function returns an int status (HUGKS_OK == 0, negative on
 *     error) and never aborts or throws across the ABI.
 *   - The caller owns all host arrays; they are read or written during the
 *     call only.  The library owns its device memory (allocated in
 *     hugks_create, released in hugks_destroy) and one HIP stream per
 *     handle.
 *   - Cell arrays are cell-major with x fastest (the iteration order of
 *     the reference's ctr(i,j) in Fortran), i.e. cell c = i + nx*j.
 *     Per-cell velocity arrays are u fastest (Fortran h(unum,vnum)),
 *     node k = iu + nv*jv.
 *   - Calls on one handle are not re-entrant; distinct handles are
 *     independent.  All calls are synchronous on return.
 */
#ifndef HUGKS_H
#define HUGKS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HUGKS_ABI_VERSION 1

/* status codes */
#define HUGKS_OK              0
#define HUGKS_E_ARG          -1  /* invalid descriptor or argument           */
#define HUGKS_E_NOMEM        -2  /* device allocation failed                 */
#define HUGKS_E_DEVICE       -3  /* HIP runtime error                        */
#define HUGKS_E_UNSUPPORTED  -4  /* size/precision without a kernel          */
#define HUGKS_E_NUMERIC      -5  /* NaN/Inf or T <= 0 met on the device      */
#define HUGKS_E_WALL         -6  /* zero outgoing wall mass coefficient      */
#define HUGKS_E_NOHIP        -7  /* no HIP device visible                    */
#define HUGKS_E_IO           -8  /* an output file could not be written      */

/* precision of f storage and per-node arithmetic */
#define HUGKS_FP64           0  /* everything FP64 (parity path)            */
#define HUGKS_MIXED_FP32     1  /* FP32 f/flux, FP64 moments and W,
                                   FP32 Hermite contractions                 */
#define HUGKS_MIXED_BF16X3   2  /* as HUGKS_MIXED_FP32 with the Hermite
                                   contractions as bf16x3 split products
                                   (operands split exactly into bf16
                                   hi+mid+lo, six products per K index on
                                   bf16 MFMA, FP32 accumulation): 2D only,
                                   MFMA node update only (nv <= 48,
                                   Nm <= 14), else HUGKS_E_UNSUPPORTED       */

/* boundary kind per side (x-lo, x-hi, y-lo, y-hi) */
#define HUGKS_BC_PERIODIC    0  /* 2d/src/boundary.f90:16-53                 */
#define HUGKS_BC_SPECULAR    1  /* 2d/src/flux.f90:551-689, 1d:188-239       */
#define HUGKS_BC_DIFFUSE     2  /* 2d/src/flux.f90:474-549, 1d:299-360       */

/* collision-time model, 2d/src/gastheory2d.f90:87-97 */
#define HUGKS_TAU_MU_OVER_RHO 0 /* case 5: muref/rho                        */
#define HUGKS_TAU_VHS         1 /* muref*2*lambda^(1-omega)/rho             */

/* CFL rule */
#define HUGKS_DT_2D_GENERAL     0 /* 2d/src/solver.f90:45-46                 */
#define HUGKS_DT_2D_WALL_NORMAL 1 /* case 5, 2d/src/solver.f90:43            */
#define HUGKS_DT_1D             2 /* 1d/src/solver.f90:38                    */

/* force_type of the reference &params namelist (2d/src/solver.f90:352-378,
 * 1d/src/solver.f90:233-242); the three 2D Hermite forms are one operator
 * (2d/tests/test_symmetric_hermite.f90:81-84 pins them equal to 1e-12) */
#define HUGKS_FORCE_HERMITE_SYMMETRIC   0 /* 2d/src/hermite_symmetric2d.f90:23  */
#define HUGKS_FORCE_HERMITE_SYMMETRIC_B 1 /* 2d/src/hermite_symmetric2d.f90:120 */
#define HUGKS_FORCE_HERMITE_LEGACY      2 /* 2d/src/gravity.f90:189; 1D cal_fv */
#define HUGKS_FORCE_FINITE_DIFFERENCE   3 /* 'finite_difference': velocity-space
                                             differences of f on the step
                                             fd_span/nv (2d/src/gravity.f90:
                                             108-187, 1d/src/gravity.f90:127-
                                             154); nv >= 3; 2D runs it in
                                             HUGKS_UPDATE_LDS only        */

/* reason an advance stopped */
#define HUGKS_STOP_NONE      0  /* step cap reached                         */
#define HUGKS_STOP_MAXTIME   1  /* simtime >= maxtime                       */
#define HUGKS_STOP_RESIDUAL  2  /* all(res < eps_res) (2D case 5)           */
#define HUGKS_STOP_TIME      3  /* simtime >= the t_stop of the call        */
#define HUGKS_STOP_STEADY    4  /* 1D steady test held steady_steps times   */
#define HUGKS_STOP_ERROR     5  /* a device flag was raised in the last step
                                   (the call returns HUGKS_E_NUMERIC or
                                   HUGKS_E_WALL)                            */

typedef struct hugks_desc {
    int32_t abi_version;      /* = HUGKS_ABI_VERSION                          */
    int32_t dim;              /* 1 or 2                                       */
    int32_t nx, ny;           /* physical cells (ny = 1 for dim = 1)          */
    int32_t nv;               /* velocity nodes per dimension                 */
    int32_t nm;               /* Hermite force order (Nm >= 0)                */
    int32_t precision;        /* HUGKS_FP64 / _MIXED_FP32 / _MIXED_BF16X3     */
    int32_t tau_model;        /* HUGKS_TAU_*                                  */
    int32_t dt_rule;          /* HUGKS_DT_*                                   */
    int32_t clamp_dt;         /* 1: dt = min(dt, maxtime - simtime) (2D)      */
    int32_t stop_on_residual; /* 1: advance stops when all(res < eps_res)     */
    int32_t bc[4];            /* x-lo, x-hi, y-lo, y-hi (1D uses bc[0..1])    */
    int32_t force_variant;    /* HUGKS_FORCE_*: the reference operator named
                                 by force_type; the library evaluates the
                                 Hermite ones with one separable operator    */
    int32_t steady_steps;     /* 1D: stop after this many consecutive steady
                                 steps (case 4, 1d/src/main.f90:109-126);
                                 0 disables the test                         */
    int32_t reserved_i[1];
    double dx, dy;            /* cell sizes (uniform grid; area = dx*dy)      */
    double pi;                /* the caller's PI constant (3.1415926535d0)    */
    double smv;               /* limiter/residual guard (2D 1e-8, 1D tiny)    */
    double innerk;            /* internal degrees of freedom K                */
    double gamma;             /* ratio of specific heats                      */
    double prandtl;           /* Prandtl number (Shakhov)                     */
    double shakhov_coef;      /* Shakhov prefactor as the caller evaluates it */
    double muref;             /* reference viscosity                          */
    double omega;             /* VHS exponent                                 */
    double cfl, maxtime;
    double eps_res;           /* residual stop threshold                      */
    double umax_dt, vmax_dt;  /* velocity bounds of the CFL rule              */
    double force_rate;        /* |phi|/velocity-gap, fixed lower bound of 1/dt */
    double tw[4];             /* wall temperature per side (diffuse walls)    */
    double tref;              /* temperature scale of the 1D steady residual  */
    double phi;               /* 1D acceleration (2D uses the arrays below)   */
    double steady_tol;        /* 1D steady test: max steady_res <= steady_tol */
    double mass_tol;          /*   and |relative mass change| <= mass_tol     */
    double fd_span;           /* velocity span lengthU of the finite-difference
                                 force (its step is fd_span/nv): umax - umin
                                 of the reference's CONFIG for Newton-Cotes
                                 grids, the node span for Gauss-Hermite ones;
                                 unused by the Hermite forces                */
    double reserved_d[1];
    const double *u;          /* [nv] u nodes                                 */
    const double *v;          /* [nv] v nodes (2D)                            */
    const double *weight;     /* [nv*nv] (2D, u fastest) or [nv] (1D)         */
    const double *prim0;      /* [(dim+2)*ncell] initial (rho,U,[V],lambda)   */
    const double *accel_cell; /* 2D: [2*ncell] (ax, ay) per cell              */
    const double *accel_vface;/* 2D: [2*(nx+1)*ny] (fn, ft), face i+(nx+1)*j  */
    const double *accel_hface;/* 2D: [2*nx*(ny+1)] (fn, ft), face i+nx*j      */
} hugks_desc;

typedef struct hugks_step_info {
    int32_t steps;            /* steps taken by this call                     */
    int32_t iter;             /* reference iteration index (starts at 1)      */
    int32_t stop_reason;      /* HUGKS_STOP_*                                 */
    int32_t device_flags;     /* bit 0 NaN/T<=0, bit 1 zero wall coefficient  */
    double simtime;
    double dt;                /* dt of the last step                          */
    double res[4];            /* residual per conserved component             */
    double steady_res[3];     /* 1D primitive-change residual                 */
    double mass_rel_change;   /* (mass - mass0) / (|mass0| + smv), mass =
                                 sum of rho * cell size; mass0 at create or
                                 at the last hugks_set_f with W              */
} hugks_step_info;

typedef struct hugks_handle hugks_handle;

/* Replaces initialize()'s allocation and distribution-function setup
 * (2d/src/initialization.f90:28-43,195-237; 1d:27-38,166-194): copies the
 * descriptor's arrays, allocates device state, sets h,b to the reduced
 * Maxwellian of prim0 (2d/src/gastheory2d.f90:6-13) and W to its
 * conserved variables (2d/src/gastheory2d.f90:28-36). */
int hugks_create(const hugks_desc *desc, hugks_handle **out);
int hugks_destroy(hugks_handle *h);

/* Message of the last failing call on h (or of the last failed create
 * when h is NULL).  Valid until the next call. */
const char *hugks_last_error(const hugks_handle *h);

/* One reference step: timestep, interpolation, evolution, update,
 * simtime += dt, iter += 1 (2d/src/main.f90:78-82,102; 1d:86-89,106-107). */
int hugks_step(hugks_handle *h, hugks_step_info *info);

/* Up to max_steps steps, with the loop exits of the reference main
 * program evaluated on the device (2d/src/main.f90:103-104; 1d:84).
 * max_steps <= 0 runs until an exit condition. */
int hugks_advance(hugks_handle *h, int32_t max_steps, hugks_step_info *info);

/* As hugks_advance, and also stops after the first step that ends with
 * simtime >= t_stop: the main program's output points (2d/src/main.f90:88-99,
 * write_macro at time2/time3/time4) without a per-step host round trip. */
int hugks_advance_until(hugks_handle *h, int32_t max_steps, double t_stop,
                        hugks_step_info *info);

/* Output fields of write_macro (2d/src/output2d.f90:156-161;
 * 1d/src/output.f90:309-321): w = conserved variables
 * [(dim+2)*ncell], q = heat flux [dim*ncell] from
 * get_heat_flux(h, b, prim(w)).  Either pointer may be NULL. */
int hugks_get_macro(hugks_handle *h, double *w, double *q);

/* As hugks_get_macro, allocating nothing: no hipMalloc, hipFree or
 * allocating library call.  This entry point has no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  q is staged in a file-scope device array of the library of
 * HUGKS_MACRO_NOALLOC_MAX doubles (8 MiB, held on each GPU it runs on until
 * the process ends): a case with dim*ncell above that is refused with
 * HUGKS_E_UNSUPPORTED (up to 524288 cells in 2D).  All calls on one GPU
 * share the array: calls that overlap on one GPU (other host threads,
 * other handles' streams) race. */
#define HUGKS_MACRO_NOALLOC_MAX (1 << 20)
int hugks_get_macro_noalloc(hugks_handle *h, double *w, double *q);

/* macro*.dat output off the step's critical path (write_macro,
 * 2d/src/output2d.f90:8-56 and 1d/src/output.f90:56-103, replaced: the
 * caller builds the columns from hugks_get_macro and hands them over).
 * Copies path, header (written as it is: the VARIABLES and ZONE lines with
 * their newlines) and nblock blocks of values (block k holds block_len[k]
 * doubles, consecutive in values) and returns; a writer thread of the
 * handle then writes the file: the header, then every block as records of
 * per_line fields in Fortran ES23.16 form with two blanks between fields,
 * each block starting a new record -- byte for byte what the reference's
 * "(11(ES23.16,2X))" (2D) and "(9(ES23.16,2X))" (1D) writes make.  Files
 * are written in call order, while the caller goes on stepping.  With h
 * NULL the file is written before the call returns.  HUGKS_E_ARG for a bad
 * argument; HUGKS_E_IO (h NULL) when the file cannot be written. */
int hugks_write_blocks_async(hugks_handle *h, const char *path, const char *header,
                             int32_t per_line, int32_t nblock, const int64_t *block_len,
                             const double *values);
/* Waits until every file queued on the handle is written; HUGKS_E_IO, with
 * the file named by hugks_last_error, when one of them could not be written
 * since the last wait.  hugks_destroy also writes what is still queued. */
int hugks_write_wait(hugks_handle *h);

/* Distribution functions as FP64, [ncell * nv^dim] each. */
int hugks_get_f(hugks_handle *h, double *hf, double *bf);
/* Overwrites h, b and (when w != NULL) W, and clears the device flags of
 * an earlier failed step.  A handle whose step raised a device flag runs
 * no further step (hugks_step/advance return the same status) until then
 * or until hugks_checkpoint_read restores another state. */
int hugks_set_f(hugks_handle *h, const double *hf, const double *bf,
                const double *w);

/* Intermediate arrays of the last step, for tests and diagnostics (the
 * reference's tests read the module globals ctr%sh, vface%flux_h, ...
 * directly, e.g. 2d/tests/test_force_and_ghosts.f90:57-58).  Written to
 * out as FP64, dense: slopes [ncell*nn], face distribution fluxes
 * [nfaces*nn] (vertical faces (nx+1)*ny, horizontal nx*(ny+1)), face
 * conserved fluxes [nfaces*4]. */
#define HUGKS_ARR_SHX    0   /* x slope of h (1D: the slope)               */
#define HUGKS_ARR_SHY    1
#define HUGKS_ARR_SBX    2
#define HUGKS_ARR_SBY    3
#define HUGKS_ARR_VFH    4   /* vertical-face flux of h (1D: face flux)    */
#define HUGKS_ARR_VFB    5
#define HUGKS_ARR_HFH    6
#define HUGKS_ARR_HFB    7
#define HUGKS_ARR_VFLUX  8   /* vertical-face conserved flux (1D: face)   */
#define HUGKS_ARR_HFLUX  9
int hugks_debug_fetch(hugks_handle *h, int32_t which, double *out);

/* Kernel schedule of the 2D step (ignored in 1D).  Every schedule computes
 * the same step; they differ in speed by case shape.
 *   flux:   HUGKS_FLUX_FUSED    one kernel per face orientation
 *           HUGKS_FLUX_SPLIT    five-kernel pipeline reading the slope arrays
 *                               (default where HUGKS_FLUX_CELLSUM does not
 *                               apply)
 *           HUGKS_FLUX_STENCIL  five-kernel pipeline forming the van Leer
 *                               slopes from each upwind cell's stencil
 *                               (needs each distribution array under 4 GiB)
 *           HUGKS_FLUX_CELLSUM  the pipeline with each face's velocity sums
 *                               added from half-space sums of its two cells
 *                               (each cell's arrays read once for the sums of
 *                               both face orientations, not once per
 *                               orientation); nv <= 64, the v nodes negative
 *                               before non-negative and weights that factor
 *                               as w[i + nv j] = wx[i] wy[j] (default where
 *                               these hold)
 *   update: HUGKS_UPDATE_LDS  Hermite contractions as LDS dot products
 *           HUGKS_UPDATE_MFMA Hermite contractions as matrix tiles
 *                             (nv <= 48, Nm <= 14)
 *           HUGKS_UPDATE_MFMA_NODE  the force and the collision terms as
 *                             matrix tiles, each node updated from its
 *                             tile accumulator (nv <= 48, Nm <= 14 and its
 *                             LDS on the device; the default where
 *                             HUGKS_UPDATE_MFMA_FACE does not apply, and the
 *                             only update of HUGKS_MIXED_BF16X3, which
 *                             refuses the others with HUGKS_E_UNSUPPORTED)
 *           HUGKS_UPDATE_MFMA_FACE  the node update that also forms the
 *                             interior faces' node fluxes: their equilibrium
 *                             part as further matrix-tile rows, the rest as
 *                             a per-cell divergence (one pass over the cells
 *                             in place of the pipeline's per-face node-flux
 *                             kernel), so no per-face node fluxes are
 *                             written.  With HUGKS_FLUX_CELLSUM or
 *                             HUGKS_FLUX_SPLIT, not with HUGKS_MIXED_BF16X3,
 *                             and distribution arrays under 4 GiB (the
 *                             default where it applies); hugks_debug_fetch
 *                             then returns the interior faces' node-flux
 *                             slots as last written. */
#define HUGKS_FLUX_FUSED   0
#define HUGKS_FLUX_SPLIT   1
#define HUGKS_FLUX_STENCIL 2
#define HUGKS_FLUX_CELLSUM 3
#define HUGKS_UPDATE_LDS   0
#define HUGKS_UPDATE_MFMA  1
#define HUGKS_UPDATE_MFMA_NODE 2
#define HUGKS_UPDATE_MFMA_FACE 3
/* A handle that belongs to a batch (hugks_batch_handle) keeps the schedule
 * and tile its batch captured at create: setting either returns
 * HUGKS_E_ARG. */
int hugks_set_schedule(hugks_handle *h, int32_t flux_kernel, int32_t update_kernel);
int hugks_get_schedule(hugks_handle *h, int32_t *flux_kernel, int32_t *update_kernel);
/* Tile parameter of HUGKS_UPDATE_MFMA_NODE and HUGKS_UPDATE_MFMA_FACE:
 * wavefronts per cell, 1 to 16
 * where a compiled kernel instance exists for the velocity grid and the
 * precision (e.g. 3, 5 and, FP32/FP64 only, 9 at nv = 33..48; 2 and 4 at
 * 17..32; 1 up to 16); 0 = the default for the handle's update: 3 for nv
 * 33..48 (5 for HUGKS_UPDATE_MFMA_FACE in FP64), 4 for 17..32, 1 up to 16;
 * hugks_create sets it.  HUGKS_E_UNSUPPORTED when no instance exists. */
int hugks_set_update_waves(hugks_handle *h, int32_t waves);
int hugks_get_update_waves(hugks_handle *h, int32_t *waves);

/* Autotuner: times every schedule and, for HUGKS_UPDATE_MFMA_NODE, every
 * wavefronts-per-cell tile with a compiled instance, on a scratch handle
 * of the descriptor's shape, and returns the fastest (cached per shape for
 * the process); *ms_per_step (may be NULL) receives its time, or -1 from
 * the cache.  hugks_tune returns the schedule only. */
int hugks_tune_tiles(const hugks_desc *d, int32_t *flux_kernel, int32_t *update_kernel,
                     int32_t *update_waves, double *ms_per_step);
int hugks_tune(const hugks_desc *d, int32_t *flux_kernel, int32_t *update_kernel,
               double *ms_per_step);

/* Batches of independent cases (a Kn / force sweep of uneven grid and
 * velocity sizes).  The reference runs each case as its own process; here
 * one step of every case, and eight steps one after the other, are
 * captured into two HIP graphs (HUGKS_BATCH_GRAPH_STEPS sets the eight);
 * an advance replays the multi-step graph, then the one-step graph for the
 * remainder, each case stopping on its own exit conditions (the launches
 * of a stopped case do nothing).  With HUGKS_BATCH_AUTOTUNE each case gets the schedule and
 * tile hugks_tune_tiles picks for its shape; a case in the grouped
 * launches (hugks_batch_grouped) keeps the grouped schedule and gets the
 * fastest wavefronts-per-cell tile of it.  hugks_batch_handle gives the case's
 * handle for hugks_get_macro and the other per-case calls (not for
 * hugks_step/advance while the batch is in use). */
#define HUGKS_BATCH_AUTOTUNE 1
typedef struct hugks_batch hugks_batch;
int hugks_batch_create(int32_t n, const hugks_desc *descs, int32_t flags,
                       hugks_batch **out);
int hugks_batch_destroy(hugks_batch *b);
const char *hugks_batch_last_error(const hugks_batch *b);
int32_t hugks_batch_size(const hugks_batch *b);
/* Number of cases stepped by grouped launches: one launch per kernel type
 * and storage type covers all small 2D cases with factoring weights and an
 * MFMA-capable grid (at least two of them; they then run the split flux
 * and the MFMA node update); the other cases' kernels are launched on their
 * own beside them.  0 when every case runs on its own. */
int32_t hugks_batch_grouped(const hugks_batch *b);
int hugks_batch_handle(hugks_batch *b, int32_t i, hugks_handle **h);
/* max_steps replays (<= 0: until every case stopped); infos[n] receives
 * each case's loop state. */
int hugks_batch_advance(hugks_batch *b, int32_t max_steps, hugks_step_info *infos);

/* Timing of a batch's grouped launches, a measurement aid: with on = 1,
 * hugks_batch_advance issues each step's launches eagerly on the batch
 * stream instead of replaying the graph (the same kernels, the same bits),
 * bracketing each grouped launch with two HIP events; the cases with their
 * own launches run untimed.  hugks_batch_kernel_times reports per grouped
 * kernel (kg_timestep, kg_slope, kg_fsum, kg_fface, kg_fnode, kg_wall,
 * kg_cell_pre, kg_update, kg_finalize) as hugks_kernel_times does. */
int hugks_batch_set_timing(hugks_batch *b, int32_t on);
int hugks_batch_kernel_times(hugks_batch *b, int32_t cap, const char **names, double *ms,
                             int32_t *calls, int32_t *count);
int hugks_batch_reset_kernel_times(hugks_batch *b);

/* Loop state (simtime, iter) of the handle. */
int hugks_get_time(hugks_handle *h, double *simtime, int32_t *iter);

/* Checkpoint / restore of the whole solver state (the reference has no
 * restart, 2d/src/initialization.f90:17-18).  The file holds the shape and
 * precision, the loop state (simtime, iter, dt, residuals, mass reference,
 * steady counter) and W, h, b exactly as stored on the device (FP32 f on
 * the mixed paths), so a handle restored from it continues bitwise as the
 * one that wrote it.  hugks_checkpoint_read needs a handle created with
 * the same dim, nx, ny, nv, Nm and precision, and with the same velocity
 * nodes, weights, accelerations, physical constants and boundary
 * conditions (compared through a hash stored in the file; loop controls
 * such as cfl and maxtime may differ); HUGKS_E_ARG names any mismatch or
 * an unreadable file. */
int hugks_checkpoint_write(hugks_handle *h, const char *path);
int hugks_checkpoint_read(hugks_handle *h, const char *path);

/* Per-kernel device time accumulated since the last reset, from HIP
 * events around each launch.  names/ms/calls have room for cap entries;
 * *count receives the number of kernels.  names point into the library. */
int hugks_kernel_times(hugks_handle *h, int32_t cap, const char **names,
                       double *ms, int32_t *calls, int32_t *count);
int hugks_reset_kernel_times(hugks_handle *h);
/* 1: bracket every launch with two HIP events recorded on the handle's
 * stream (no synchronisation per launch; the elapsed times are collected
 * by hugks_kernel_times, or after 4096 launches); 0: off (default). */
int hugks_set_timing(hugks_handle *h, int32_t on);

/* Library/device information string (gfx target, CU count). */
const char *hugks_device_info(void);

#ifdef __cplusplus
}
#endif
#endif /* HUGKS_H */
