/*
 * flowperm.h -- C ABI of libflowperm.so, the MI355X lattice-Boltzmann
 * permeability solver.
 *
 * The reference has no library boundary: its solvers are executables driven
 * through subprocess.run (scripts/solve_permeability3d.py:113-146,
 * scripts/solve_permeability.py:29-46), reading a mask file and appending one
 * CSV row, with exit codes 0 / 3 / 2 (LMB3d/lbm3d-perm.cpp:140-144, :630;
 * LMB2d/main.cpp:426).  This ABI moves that process boundary to a host ->
 * device boundary: masks and results cross as plain memory, and every status
 * maps 1:1 onto the reference exit codes (fp_exit_code).
 *
 * Conventions
 *  - Masks are uint8, 1 = solid, 0 = pore.  3D masks are [nz][ny][nx] with x
 *    fastest (the .raw payload, LMB3d/lbm3d-perm.cpp:156); 2D masks are
 *    [ly][lx], row = solver y, column = solver x (scripts/structure_io.py:44-54).
 *  - The caller owns every input and output buffer; the library owns device
 *    memory.  No C++ exception crosses the ABI.
 *  - A context is used by one thread at a time; different contexts are
 *    independent (there is no global state).
 *  - Statuses: FP_OK / FP_NOT_CONVERGED are the reference's exit codes 0 / 3;
 *    every FP_ERR_* is its exit code 2 (a message is kept in fp_last_error).
 */
#ifndef FLOWPERM_H
#define FLOWPERM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FP_ABI_VERSION 10

enum fp_status {
    FP_OK              = 0,   /* all directions converged            (exit 0) */
    FP_PENDING         = 1,   /* server request not finished yet (no exit code) */
    FP_NOT_CONVERGED   = 3,   /* a direction hit max_steps; row kept (exit 3) */
    FP_ERR_BAD_DIMS    = 10,  /* read_structure: bad dimensions      (exit 2) */
    FP_ERR_NON_BINARY  = 11,  /* read_structure: non-binary voxel    (exit 2) */
    FP_ERR_NO_PORE     = 12,  /* read_structure: no pore space       (exit 2) */
    FP_ERR_NONFINITE   = 13,  /* solve(): non-finite flux            (exit 2) */
    FP_ERR_OOM         = 14,  /* device memory exhausted             (exit 2) */
    FP_ERR_DEVICE      = 15,  /* HIP runtime failure                 (exit 2) */
    FP_ERR_ARG         = 16,  /* invalid argument                    (exit 2) */
    FP_ERR_NOT_FOUND   = 17   /* unknown job id                      (exit 2) */
};

enum fp_precision {
    FP_F64        = 0,  /* populations in double: the reference's arithmetic (3D),
                           the 2D scheme evaluated in double (2D)               */
    FP_F32        = 1,  /* populations stored as float f - w; flux sums, the
                           convergence test and K stay in double              */
    FP_F32_REF2D  = 2   /* 2D only: unshifted float, the arithmetic of
                           LMB2d/lbm.cpp:516-585 operation for operation       */
};

/* Process exit code of the reference for a status: 0, 3 or 2. */
int fp_exit_code(int status);

/* ------------------------------------------------------------------------ */
/* contexts                                                                  */
/* ------------------------------------------------------------------------ */
typedef struct fp_ctx fp_ctx;

/* Open a context on HIP device `device` (its own stream).  Replaces the
 * process start of lbm3d-perm / lbm2d-perm (LMB3d/lbm3d-perm.cpp:478-481). */
fp_ctx *fp_create(int device, int *status);
void    fp_destroy(fp_ctx *ctx);
/* Message of the last failure on this context (stderr text of die(),
 * LMB3d/lbm3d-perm.cpp:140-144), or "" */
const char *fp_last_error(fp_ctx *ctx);
int     fp_abi_version(void);

/* ------------------------------------------------------------------------ */
/* 3D: D3Q19                                                                 */
/* ------------------------------------------------------------------------ */
/* Config + flags of LMB3d/lbm3d-perm.cpp:118-130 and :487-505. */
typedef struct fp_config3d {
    double  ff;             /* --ff: body force = 2.5e-7 * ff         (1)      */
    double  eps;            /* --eps: drift tolerance relative to |q| (1e-5)   */
    int64_t lag;            /* --lag: steps between checks            (2000)   */
    int64_t min_steps;      /* --min-steps                            (4000)   */
    int64_t max_steps;      /* --max-steps                            (200000) */
    int64_t avg_steps;      /* --avg-steps                            (2000)   */
    int32_t hold;           /* --hold                                 (2)      */
    int32_t validate;       /* --validate: 4th solve along (1,1,1)/sqrt(3)     */
    int32_t precision;      /* enum fp_precision                      (FP_F64) */
    int32_t concurrent;     /* step the force directions together     (0)      */
} fp_config3d;

/* Fill with the reference defaults (LMB3d/lbm3d-perm.cpp:118-130). */
void fp_config3d_default(fp_config3d *cfg);

/* One forced solve: the Solve struct of LMB3d/lbm3d-perm.cpp:132-138. */
typedef struct fp_solve_out {
    double  q[3];           /* Darcy flux (superficial velocity) per axis     */
    double  se[3];          /* standard error over the averaging phase        */
    int64_t steps;
    int32_t converged;
    int32_t status;         /* FP_OK or FP_ERR_NONFINITE                      */
    double  fieldres;       /* 2D only: field residual (LMB2d/lbm.cpp:169-182) */
    /* The last first-phase convergence check, the values the reference's
     * progress line prints (LMB3d/lbm3d-perm.cpp:353-357 "# fx S q=(..)
     * drift/tol=R hold=H"; LMB2d/main.cpp:167-172 "# fx S qx=.. qy=..
     * dx/tol=.. dy/tol=.. hold=H").  For a run that reached max_steps,
     * check_ratio says how far from converged it stopped. */
    int64_t check_step;     /* step of that check; 0 = no check ran           */
    int64_t checks;         /* checks in the first phase so far               */
    double  check_q[3];     /* 3D: the flux q; 2D: the block means qx, qy     */
    double  check_ratio[2]; /* 3D: drift/tol in [0]; 2D: dx/tol, dy/tol       */
    int32_t check_hold;     /* hold after that check                          */
    int32_t pad;
} fp_solve_out;

/* Progress of fp_solve3d / fp_solve2d / fp_solve_dir3d / fp_solve_dir2d on
 * this context: after every first-phase convergence check of a run, fn(user,
 * run, out) with the run's index in the call (0 fx, 1 fy, 2 fz or 2D val,
 * 3 3D val) and its check_* fields, steps (the run's step count) and
 * converged filled.  It replaces the per-check progress lines of the
 * reference (LMB3d/lbm3d-perm.cpp:353-357, LMB2d/main.cpp:167-172), which the
 * executables print from it unless --quiet.  While a callback is set, the
 * runs are stepped in chunks no longer than the check interval (lag), so
 * every check is reported (with lag 1, every other one).  NULL turns it off.
 * The callback runs on the calling thread, between chunks. */
typedef void (*fp_progress_fn)(void *user, int run, const fp_solve_out *out);
int     fp_set_progress(fp_ctx *ctx, fp_progress_fn fn, void *user);

/* Everything lbm3d-perm writes to its CSV row (LMB3d/lbm3d-perm.cpp:574-615)
 * plus the --validate tensor test (:555-572). */
typedef struct fp_result3d {
    int32_t nx, ny, nz, pad0;
    int64_t n_pore;
    double  porosity, force, mu;
    fp_solve_out run[3];    /* forcing along +x, +y, +z                       */
    double  K[3][3];        /* K[flux component][force axis]                  */
    double  k_xy, k_xz, k_yz;
    double  recip_xy, recip_xz, recip_yz, recip_resid;
    double  k[3];           /* principal values, descending                   */
    double  k_mean, fa;
    double  v[3][3];        /* v[component][eigen index] (cyclic Jacobi, :401-443) */
    int32_t validated, pad1;
    fp_solve_out val;       /* the (1,1,1) solve                              */
    double  val_pred[3];
    double  val_err;
    double  seconds;        /* wall time of the solves                        */
    int64_t fluid_updates;  /* fluid-cell updates performed                   */
} fp_result3d;

/* Check a mask as read_structure does (LMB3d/lbm3d-perm.cpp:182-192):
 * FP_OK, FP_ERR_BAD_DIMS, FP_ERR_NON_BINARY or FP_ERR_NO_PORE.  n_pore may be
 * NULL. */
int fp_check_mask3d(const uint8_t *solid, int nx, int ny, int nz, int64_t *n_pore);

/* The whole lbm3d-perm invocation minus file I/O (LMB3d/lbm3d-perm.cpp:507-630):
 * three solves along +x, +y, +z (+ the tensor test when cfg->validate), K,
 * reciprocity, eigen-decomposition and FA.  vel, when not NULL, receives the
 * export_velocity payload of each direction (:305-317): 3 runs x NN x 3
 * floats (U, V, W per node, x fastest), i.e. what PREFIX.f{x,y,z}.vel hold
 * after their 16-byte header.  Returns FP_OK, FP_NOT_CONVERGED or an error. */
int fp_solve3d(fp_ctx *ctx, const uint8_t *solid, int nx, int ny, int nz,
               const fp_config3d *cfg, fp_result3d *out, float *vel);

/* One forcing direction: solve() of LMB3d/lbm3d-perm.cpp:320-398 for the body
 * force F (lattice units).  vel: NULL or NN*3 floats. */
int fp_solve_dir3d(fp_ctx *ctx, const uint8_t *solid, int nx, int ny, int nz,
                   const fp_config3d *cfg, const double F[3], fp_solve_out *out, float *vel);

/* fp_solve_dir3d without device allocation (no counterpart in the reference;
 * for callers that cannot allocate device memory: code that must not call
 * hipMalloc, tight serving loops, harnesses that forbid allocation).  The same
 * forced solve, bit for bit, with no velocity export: every device buffer of
 * it -- tables, populations, state, work lists, flux partials -- lies in one
 * array of FP_NOALLOC_SCRATCH_BYTES that the library holds on each device it
 * runs on until the process ends, and no hipMalloc, hipFree or allocating
 * library call is made.  A lattice whose buffers exceed the array is rejected
 * with FP_ERR_ARG (the bytes it needs in fp_last_error): 64^3 in FP64 and
 * 80^3 in FP32 fit at any porosity.  Every call on a device uses that device's
 * one array and leaves its contents behind; calls on one device are serialised
 * by the library (two at once would race on the array). */
#define FP_NOALLOC_SCRATCH_BYTES 134217728
int fp_solve_dir3d_noalloc(fp_ctx *ctx, const uint8_t *solid, int nx, int ny, int nz,
                           const fp_config3d *cfg, const double F[3], fp_solve_out *out);

/* Self-check of the neighbour tables the step kernels use (no reference
 * counterpart; host only, no device work): for every fluid cell and lattice
 * direction, the neighbour index the kernels derive from the per-segment link
 * records against a direct lookup on the dense periodic lattice.  dim 3 =
 * D3Q19 ([nz][ny][nx]), dim 2 = D2Q9 ([ny][nx], nz ignored).  *mismatches = 0
 * when every link is right. */
int fp_verify_links(const uint8_t *solid, int dim, int nx, int ny, int nz, int64_t *mismatches);

/* Symmetric 3x3 eigen-decomposition exactly as jacobi3 (:401-443). */
void fp_jacobi3(const double A[9], double val[3], double vec[9]);

/* ------------------------------------------------------------------------ */
/* 2D: D2Q9                                                                  */
/* ------------------------------------------------------------------------ */
/* Config + flags of LMB2d/main.cpp:64-79 and :286-308; L0 of LMB2d/lbm.h:27. */
typedef struct fp_config2d {
    double  ff;             /* (1)       */
    double  eps;            /* (1e-3)    */
    double  qfloor;         /* (1e-2)    */
    double  ksigma;         /* (3)       */
    double  L0;             /* flux normalisation length (4 at 256x256; the
                               reference scales it with the lattice, lbm.h:24-27) */
    double  validate_deg;   /* --validate DEG; < 0 = off                    */
    int64_t lag;            /* (5000)    */
    int64_t min_steps;      /* (10000)   */
    int64_t max_steps;      /* (500000)  */
    int64_t avg_steps;      /* (20000)   */
    int64_t avg_sample;     /* (50)      */
    int32_t hold;           /* (3)       */
    int32_t precision;      /* enum fp_precision (FP_F32_REF2D = the reference) */
    int32_t concurrent;
    int32_t pad;
} fp_config2d;

void fp_config2d_default(fp_config2d *cfg);

/* field_residual of LMB2d/lbm.cpp:169-182 (the fieldres_* columns): the L2
 * change between two velocity fields, both [ly][lx] rows of (U, V) floats,
 * relative to the first: sqrt(sum |cur - snap|^2 / sum |cur|^2), rounded to
 * float, or 10 when cur is zero everywhere.  Host only. */
double fp_field_residual2d(int lx, int ly, const float *cur, const float *snap);

/* The convergence rules of solve() alone, without a lattice (host only, no
 * device work): the state machine the device advances after every step,
 * driven by a given flux sequence.  q holds nsteps x 3 (2D: x 2) values; step
 * t (1-based) sees q[(t-1)*3 ..] as the flux volumeavg would return after it
 * when the rules ask for one that step, and nothing otherwise.  Fills out
 * (q, se, steps, converged, status) as the solve would and returns FP_OK,
 * FP_NOT_CONVERGED or FP_ERR_NONFINITE; FP_ERR_ARG when the sequence ends
 * before the solve does.  3D: phase 1 and 2 of LMB3d/lbm3d-perm.cpp:320-398;
 * 2D: the block-mean / k-sigma test and averaging of LMB2d/main.cpp:99-235. */
int fp_replay_convergence3d(const fp_config3d *cfg, const double *q, int64_t nsteps,
                            fp_solve_out *out);
int fp_replay_convergence2d(const fp_config2d *cfg, const double *q, int64_t nsteps,
                            fp_solve_out *out);

/* Everything lbm2d-perm writes to its CSV row (LMB2d/main.cpp:382-411). */
typedef struct fp_result2d {
    int32_t lx, ly;
    double  porosity, force, mu;
    fp_solve_out run[2];    /* forcing along +x, +y (q[2], se[2] unused)     */
    double  k_xx, k_yx, k_xy, k_yy, k_off, recip_resid;
    double  k1, k2, k_mean, anisotropy, theta_deg, alpha_x_deg, alpha_y_deg;
    int32_t validated, pad1;
    fp_solve_out val;
    double  val_pred[2];
    double  val_err;
    double  seconds;
    int64_t fluid_updates;
} fp_result2d;

/* The whole lbm2d-perm invocation minus file I/O (LMB2d/main.cpp:310-426).
 * solid is [ly][lx].  vel: NULL or 2 runs x (ly*lx) x 2 floats ([y][x] rows
 * of (U, V), what exportvelocity, LMB2d/lbm.cpp:587-594, prints). */
int fp_solve2d(fp_ctx *ctx, const uint8_t *solid, int lx, int ly,
               const fp_config2d *cfg, fp_result2d *out, float *vel);

/* One forcing direction: solve() of LMB2d/main.cpp:99-235. */
int fp_solve_dir2d(fp_ctx *ctx, const uint8_t *solid, int lx, int ly,
                   const fp_config2d *cfg, const double F[2], fp_solve_out *out, float *vel);

/* ------------------------------------------------------------------------ */
/* batched solving (no reference counterpart: the reference runs one process
 * per structure, scripts/solve_permeability3d.py:250).  Every job of a batch
 * is one forced solve; all live jobs advance together, each retires on the
 * device when its own convergence test passes.                              */
/* ------------------------------------------------------------------------ */
typedef struct fp_batch fp_batch;

/* lattice: 3 (D3Q19) or 2 (D2Q9); precision: enum fp_precision */
fp_batch *fp_batch_create(fp_ctx *ctx, int lattice, int precision, int *status);
void      fp_batch_destroy(fp_batch *b);
/* add one forced solve of a 3D mask with cfg's convergence settings; returns
 * a job id >= 0, or -status.  Masks with identical bytes share one device
 * geometry. */
int64_t   fp_batch_add3d(fp_batch *b, const uint8_t *solid, int nx, int ny, int nz,
                         const fp_config3d *cfg, const double F[3]);
int64_t   fp_batch_add2d(fp_batch *b, const uint8_t *solid, int lx, int ly,
                         const fp_config2d *cfg, const double F[2]);
/* run every job to completion */
int       fp_batch_run(fp_batch *b);
int       fp_batch_result(fp_batch *b, int64_t id, fp_solve_out *out);
/* steps launched, lattice-cell updates performed, seconds on the device */
void      fp_batch_stats(fp_batch *b, int64_t *steps, int64_t *fluid_updates, double *seconds);
/* steps per hipGraph chunk (even; default 200) */
void      fp_batch_set_chunk(fp_batch *b, int steps);

/* ------------------------------------------------------------------------ */
/* measurement (no reference counterpart: the reference timed its solver by
 * hand, NOTES.md:1193-1205).  Steps one lattice at rest under force F for
 * `warmup` then `steps` steps (even counts), with a HIP event pair around
 * every kernel on the context's stream.                                     */
/* ------------------------------------------------------------------------ */
typedef struct fp_bench_out {
    int64_t n_fluid;          /* fluid cells (one update each per step)     */
    int64_t n_cells;          /* all lattice cells                          */
    int64_t steps;            /* timed steps                                */
    double  total_ms;         /* event time of the timed steps, replayed as */
                              /* hipGraph chunks as solves run them         */
    double  local_ms;         /* average duration: local (odd) step kernel  */
    double  nbr_ms;           /* average duration: neighbour (even) kernel  */
    double  monitor_ms;       /* average duration: flux reduction and       */
                              /* convergence update after a step (k_finish) */
    double  bytes_per_update; /* algorithmic bytes per fluid update: 2 Q s  */
    double  eager_total_ms;   /* the same steps launched one by one with an */
                              /* event pair around every kernel (the pass   */
                              /* the per-kernel averages come from)         */
    int32_t deferred;         /* 1: total_ms ran with the finish of each    */
                              /* step deferred into the next step kernel    */
                              /* (lattices of at most 1024 work items per   */
                              /* step kernel), one finish launch per chunk; */
                              /* 0: finish launches, after every step or    */
                              /* only where the chunk's finish plan puts    */
                              /* them (a plan is taken when it keeps at     */
                              /* most one finish per 5 steps)               */
    int32_t pad;
} fp_bench_out;

/* Launch shapes of the step kernels; the autotuner picks them per device.  A new
   context starts with local_cpt 1, nbr_waves 4, chunk_steps 200, nontemporal 1,
   local_bpc 0, nbr_bpc 0, nbr_mode 2, nbr_segs 4, local_wgs 3, nbr_wgs 5 (the
   shapes tuned for D3Q19 FP32 at 256^3). */
typedef struct fp_tuning {
    int32_t local_cpt;        /* cells per thread, local kernel: 1, 2 or 4            */
    int32_t nbr_waves;        /* row segments (wavefronts) per workgroup, neighbour
                                 kernel: 1..8                                         */
    int32_t chunk_steps;      /* steps per hipGraph replay (even)                     */
    int32_t nontemporal;      /* 1: local-kernel loads/stores do not allocate in cache */
    int32_t local_bpc;        /* workgroups launched per CU, local kernel: 0..32
                                 (0: one workgroup per work item)                     */
    int32_t nbr_bpc;          /* the same for the neighbour kernel                    */
    int32_t nbr_mode;         /* neighbour step: 0 = one wavefront per 64-cell row
                                 segment, 1 = 64 fluid cells per wavefront (a 24-byte
                                 per-cell neighbour record; faster below porosity
                                 ~0.7), 2 = chosen per structure by porosity (default) */
    int32_t nbr_segs;         /* of the nbr_waves segments, how many lie side by side
                                 along x in one row: 1, 2, 4 or 8 (at most nbr_waves,
                                 and lowered to a divisor of the row's 64-cell
                                 segments; the others stack rows along z, along y in
                                 2D); 0 = 4 (default)                                 */
    int32_t local_wgs;        /* resident workgroups per CU the local kernel is held
                                 to, through LDS it does not use: 1..6 on a 160 KiB
                                 CU (a cap no LDS size reaches is not applied); 0 =
                                 as many as its registers allow.  Applied to D3Q19
                                 chunks that run a finish plan (most steps without a
                                 flux sum); chunks with a finish after every step or
                                 the deferred finish, and D2Q9, launch uncapped.  Fewer streams in
                                 flight move more bytes: 256^3 FP32 local step 0.235
                                 against 0.245 ms at 3.                              */
    int32_t nbr_wgs;          /* the same for the neighbour kernels (0.348 against
                                 0.353 ms at 5)                                      */
} fp_tuning;

/* Set / read the launch shapes a context uses for its solves and batches. */
int  fp_set_tuning(fp_ctx *ctx, const fp_tuning *t);
void fp_get_tuning(fp_ctx *ctx, fp_tuning *t);

/* Self-check of a batch's flux-partial layout (no reference counterpart; host
 * only, no device work): builds the tables a batch would for the forcing runs
 * of n structures (3 per 3D mask, 2 per 2D one; dims[3i..3i+2] = nx, ny, nz of
 * mask i, nz ignored in 2D), with the neighbour-step forms chosen as the batch
 * chooses them under tune (NULL: the defaults), and checks that every
 * flux-partial slot a step kernel writes lies in its array and in its own
 * instance's range, that the instances' ranges partition each array and that
 * each instance's reduction items tile its range.  *problems = 0 when all
 * hold. */
int fp_verify_batch_tables(const uint8_t *const *solids, const int32_t *dims, int n, int lattice,
                           int precision, const fp_tuning *tune, int64_t *problems);

/* Self-check of the sparse finish schedule (no reference counterpart; host
 * only, no device work).  A batch's chunks launch the per-step finish (flux
 * sum and convergence update) only after steps on which some instance's rules
 * can act, and the next finish counts the quiet steps in between.  This runs
 * the convergence rules of n instances -- c3 (3D) or c2 (2D) configurations,
 * exactly one non-NULL; q holds nsteps x 3 flux values per instance, step t
 * (1-based) of instance i seeing q[(i*nsteps + t-1)*3 ..]; instance i joins
 * at batch step start[i] -- once with a finish after every step and once with
 * the finish schedules of chunks of `chunk` steps (even) replayed `replays`
 * times, and counts in *problems every step whose outputs (phase, flux sum,
 * velocity write) and every finish whose state differ.  *finishes: finishes
 * of the schedules.  plant = 1 replaces each schedule by one finish at the
 * chunk's end, a wrong schedule the check must catch. */
int fp_verify_finish_plan(const fp_config3d *c3, const fp_config2d *c2, int n, const double *q,
                          int64_t nsteps, const int64_t *start, int chunk, int replays, int plant,
                          int64_t *problems, int64_t *finishes);

/* Steps one resident instance: `warmup` then `steps` steps launched one by one
 * with events around every kernel (the per-kernel averages), then `steps` more
 * as hipGraph chunk replays (total_ms).  tune: NULL = the context's shapes.
 * flux_every_step: 0 = convergence checks at the default lag (2000 steps), as
 * in the first phase of a solve; 1 = the exact flux sum on every step, as in
 * the averaging phase. */
int fp_bench_steps(fp_ctx *ctx, const uint8_t *solid, int nx, int ny, int nz, int lattice,
                   int precision, const double F[3], int warmup, int steps,
                   const fp_tuning *tune, int flux_every_step, fp_bench_out *out);

/* Autotuner: times every launch shape of the two step kernels on the given
 * lattice (`steps` steps each), stores the fastest in *best and makes it the
 * context's tuning.  When not NULL, local_ms / nbr_ms receive the average
 * duration of every local / neighbour shape tried (FP_TUNE_NLOCAL /
 * FP_TUNE_NNBR entries, in the order of fp_tune_shape). */
#define FP_TUNE_NLOCAL 24
#define FP_TUNE_NNBR   16
int fp_autotune(fp_ctx *ctx, const uint8_t *solid, int nx, int ny, int nz, int lattice,
                int precision, int steps, fp_tuning *best, double *local_ms, double *nbr_ms);
/* The i-th shape tried by fp_autotune: kind 0 = local (cpt, bpc, nontemporal in
 * a, b, c), kind 1 = neighbour (waves, bpc, nbr_mode in a, b, c), kind 2 = the
 * same neighbour shape's nbr_segs (in a), kind 3 = the local shape's local_wgs,
 * kind 4 = the neighbour shape's nbr_wgs (in a). */
void fp_tune_shape(int kind, int i, int *a, int *b, int *c);

/* ------------------------------------------------------------------------ */
/* request server (no reference counterpart: the reference's drivers run one
 * solver process per structure from a process pool,
 * scripts/solve_permeability3d.py:250, and a structure's three forcing runs
 * one after another inside it, LMB3d/lbm3d-perm.cpp:518-521).
 *
 * A request is one whole lbm3d-perm / lbm2d-perm invocation: its 3 (2D: 2)
 * forcing runs, plus the --validate run when asked, become instances of one
 * continuously batched stepper.  Requests wait until the device memory they
 * need (segment table + population arrays) fits under the budget, oldest
 * first; a request that fits may be admitted ahead of an older one that does
 * not, at most max_bypass times per waiting request.  Admitted requests join
 * the running batch at the next chunk boundary, each
 * instance retires on the device when its own convergence test passes, and a
 * request's memory is returned as soon as its last instance retires.
 * A request's result does not depend on what else is resident (every instance
 * is stepped by the same kernels on its own arrays, and flux sums are exact).
 *
 * Threads: fp_server_submit* and fp_server_poll* may be called from any
 * thread while one thread drives fp_server_step.                           */
/* ------------------------------------------------------------------------ */
typedef struct fp_server fp_server;

typedef struct fp_server_config {
    int32_t lattice;          /* 3 (D3Q19) or 2 (D2Q9)                          */
    int32_t precision;        /* enum fp_precision, for every request            */
    int64_t budget_bytes;     /* device memory for resident requests; 0 = 90% of
                                 the free memory when the server is created      */
    int32_t chunk_steps;      /* steps between admissions (even); 0 = context's  */
    int32_t max_bypass;       /* later requests that may be admitted ahead of a
                                 waiting one that does not fit yet, before it
                                 blocks the queue; 0 = 16, < 0 = strict arrival
                                 order                                          */
} fp_server_config;

typedef struct fp_request_info {
    double  t_submit;         /* fp_now() when submitted                         */
    double  t_admit;          /* when its memory was granted (0 while queued)    */
    double  t_done;           /* when its last instance retired (0 while pending) */
    int64_t device_bytes;     /* footprint charged against the budget            */
    int64_t fluid_updates;    /* fluid-cell updates of its instances             */
} fp_request_info;

typedef struct fp_server_stats {
    int64_t submitted, admitted, completed, rejected;
    int64_t queued;           /* waiting for memory                              */
    int64_t resident;         /* admitted, not finished                          */
    int64_t live_instances;   /* forcing runs still stepping                     */
    int64_t chunks, steps, fluid_updates;
    int64_t bytes_in_use, bytes_peak, budget_bytes;
    double  busy_seconds;     /* wall time spent stepping on the device          */
} fp_server_stats;

/* seconds on the library's monotonic clock (the fp_request_info times) */
double     fp_now(void);
fp_server *fp_server_create(fp_ctx *ctx, const fp_server_config *cfg, int *status);
void       fp_server_destroy(fp_server *s);
/* Queue one lbm3d-perm invocation (cfg->precision must be the server's;
 * cfg->concurrent is implied).  Returns a request id >= 0, or -status: mask
 * errors as fp_check_mask3d, FP_ERR_OOM when the request alone exceeds the
 * budget. */
int64_t    fp_server_submit3d(fp_server *s, const uint8_t *solid, int nx, int ny, int nz,
                              const fp_config3d *cfg);
int64_t    fp_server_submit2d(fp_server *s, const uint8_t *solid, int lx, int ly,
                              const fp_config2d *cfg);
/* Admit what fits, step every resident instance for one chunk, retire and
 * assemble finished requests.  *pending (may be NULL) = requests queued or
 * resident afterwards; 0 means the server is idle.  Returns FP_OK or a device
 * error (every resident request then fails with it). */
int        fp_server_step(fp_server *s, int64_t *pending);
/* FP_PENDING while queued or stepping; otherwise the request's final status
 * (as fp_solve3d / fp_solve2d would return) with *out filled, after which the
 * id is forgotten (FP_ERR_NOT_FOUND).  info may be NULL. */
int        fp_server_poll3d(fp_server *s, int64_t id, fp_result3d *out, fp_request_info *info);
int        fp_server_poll2d(fp_server *s, int64_t id, fp_result2d *out, fp_request_info *info);
void       fp_server_get_stats(fp_server *s, fp_server_stats *out);

/* Tests of the error paths (no reference counterpart): the next n table
 * rebuilds of any batch or server in the process fail as a device error
 * would, before they touch the device.  0 clears it. */
void fp_test_fail_rebuilds(int n);

/* Plain device-to-device copy of `bytes` (read + write), `iters` times with
 * each of four copy kernels (a persistent grid; one workgroup per 8 KiB; the
 * same with non-temporal loads and stores; one workgroup per 4 KiB,
 * non-temporal): *ms_per_copy is the fastest, the measured bandwidth the step
 * kernels are compared with. */
int fp_bench_copy(fp_ctx *ctx, int64_t bytes, int iters, double *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif /* FLOWPERM_H */
