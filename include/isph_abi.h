/*
 * isph_abi.h -- C ABI of the MI355X-native ISPH per-step particle pipeline.
 *
 * The reference program drives its per-step pipeline through argument-less
 * Fortran module subroutines that read and write module globals
 * (main.f90:199-272; state owned by initialize.f90:21-83).  This header
 * replaces those subroutines with one entry point each, operating on an
 * opaque device-resident context.  A Fortran driver binds them with
 * iso_c_binding (beach-slope_amd/fortran/isph_gpu.f90); Python binds them
 * with ctypes (beach-slope_amd/__init__.py).
 *
 * Conventions
 *   - Every function returns 0 on success or a negative ISPH_E* code; the
 *     message of the last failure is returned by isph_last_error().
 *   - Host arrays passed in are copied, never retained.  Particle state is
 *     device resident between calls; only downloads synchronise.
 *   - Particles are held in cell-sorted order: the index of a particle is the
 *     reference's PPE row / matid minus one after `resetid` (isph.f90:88-113).
 *   - A context owns one HIP stream and is not thread safe.
 *   - Cells are numbered (ci, cj) with ci = 0..celly-1 counted from the top
 *     of the domain and cj = 0..cellx-1 from the left; the reference's
 *     dpcell(i,j) is (ci, cj) = (i-1, j-1).  Cell key = cj*celly + ci, the
 *     column-major order that resetid walks (isph.f90:97-103).
 */
#ifndef ISPH_ABI_H
#define ISPH_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ISPH_OK = 0,
  ISPH_EHIP = -1,       /* HIP runtime error                                  */
  ISPH_EARG = -2,       /* bad argument (size, null pointer, particle outside) */
  ISPH_ECAP = -3,       /* capacity overflow (particles, neighbours, cells)   */
  ISPH_ENONFINITE = -4, /* zero / non-finite PPE diagonal or residual         */
  ISPH_ENOCONV = -5,    /* PPE solver did not reach its tolerance             */
  ISPH_ESTATE = -6      /* call out of order (e.g. solve before assembly)     */
};

/* Particle type ids (particle.f90:11): 1 fixed wall, 2 ghost wall, 3 fluid. */
enum { ISPH_TID_FIXED = 1, ISPH_TID_GHOST = 2, ISPH_TID_FLUID = 3 };

/* Flag bits carried per particle (the logicals of particle.f90:42-43). */
enum { ISPH_F_BUFFER = 1, ISPH_F_FREE = 2, ISPH_F_VICINITY = 4 };
/* A wall particle (tid 1 or 2) of the wavemaker paddle; it has no
 * counterpart in the reference, whose walls never move. */
enum { ISPH_F_PADDLE = 8 };

/* Maximum neighbours per particle (the reference's per-slot capacity
 * `fplist`, memory.f90:25).  Exceeding it is ISPH_ECAP. */
#define ISPH_NBR_CAP 64

/* Upper bound on the scalar sub-steps per step (solsteps of timestep,
 * integrator.f90:943-949); a larger count is reported as ISPH_ENONFINITE. */
#define ISPH_MAX_SOLSTEPS 100000

/* PPE solve forms.
 *   ISPH_PPE_FP64        BiCGStab in FP64, iterates identical to the CPU
 *                        restatement (the parity path).
 *   ISPH_PPE_FP32_REFINE BiCGStab right-preconditioned by a cell-aggregation
 *                        multigrid V-cycle (FP16 level-0 matrix, FP32
 *                        vectors); FP32 search directions; FP64 products of
 *                        the matrix with the directions, solution, residual
 *                        and dot products; every solve starts from x = 0;
 *                        a convergence claim is checked against the FP64
 *                        residual b - A x and the iteration restarts from it
 *                        when the check fails (iterative refinement). */
enum { ISPH_PPE_FP64 = 0, ISPH_PPE_FP32_REFINE = 1 };

/*
 * Run parameters: the module variables of initialize.f90:21-83 plus the
 * constants hard-coded in porous.f90:111, turbulence.f90:16 and the derived
 * values of setup.f90:10-150.  All lengths in metres, SI units.
 */
typedef struct isph_params {
  int32_t cellx, celly;   /* cell grid incl. one halo ring (setup.f90:51-52)   */
  int32_t fixed_dt;       /* 1: dt stays fixed; 0: adaptive (integrator.f90:927-937) */
  int32_t precision;      /* ISPH_PPE_FP64 or ISPH_PPE_FP32_REFINE             */
  int32_t max_iter;       /* BiCGStab iteration cap per solve                  */
  int32_t max_outer;      /* restarts cap of the precision form                */
  double h;               /* smoothing length h1 (setup.f90:47)                */
  double dl;              /* lattice spacing dl (setup.f90:53)                 */
  double lam;             /* 1e-3*h^2 (setup.f90:82)                           */
  double rho;             /* reference density (initialize.f90:23)             */
  double rel_den;         /* rhomax/rho (initialize.f90:27)                    */
  double rhomin;          /* (initialize.f90:24)                               */
  double g;               /* gravity, negative (initialize.f90:23)             */
  double mu;              /* (initialize.f90:39)                               */
  double por;             /* porosity of the porous layer (initialize.f90:26)  */
  double d50;             /* grain size (porous.f90:111)                       */
  double lamk;            /* Darcy constant (porous.f90:111)                   */
  double csh;             /* shifting coefficient (initialize.f90:24)          */
  double maxshift;        /* shift cap factor (initialize.f90:40)              */
  double fs_thresh;       /* lamfs*maxdivr (initialize.f90:45-46)              */
  double sig1;            /* time-step coefficient (initialize.f90:39)         */
  double tschmidt;        /* turbulent Schmidt number (initialize.f90:25)      */
  double xlcutoff;        /* outflow cut of cellshiftalt (integrator.f90:61)   */
  double ytcutoff;
  double buf_xmin;        /* buffer release window of comp_pos (integrator.f90:867-868) */
  double buf_xmax;
  double dt;              /* initial (or fixed) time step                      */
  double tol;             /* PPE relative residual target ||b-Ax||/||b||       */
  double inner_tol;       /* precision form: the recursive residual stops at
                             inner_tol * tol (0: 0.05) before the FP64 check  */
} isph_params;

/* Statistics of the last PPE solve. */
typedef struct isph_solve_stats {
  int32_t iterations;     /* BiCGStab iterations (summed over restarts)        */
  int32_t outer;          /* restarts of the precision form (0 for FP64 path)  */
  int32_t converged;      /* 1 if the target was met                           */
  int32_t breakdown;      /* 1 if rho or omega became zero / non-finite        */
  double rel_residual;    /* final ||b-Ax||/||b|| (recursive for FP64 path)    */
  double bnorm;           /* ||b||                                             */
  double nnz;             /* stored PPE entries (sum of row lengths)           */
} isph_solve_stats;

/* Time-stepping scalars (module variables t, dt, dtsol, solsteps, umax, numax). */
typedef struct isph_scalars {
  double t, dt, dtsol, umax, numax;
  int32_t solsteps;
  int32_t n;              /* live particle count                               */
} isph_scalars;

typedef struct isph_ctx isph_ctx;

/* ---- lifecycle (replaces init/memalloc/matrixid, setup.f90:10, memory.f90:10-145) ---- */
int isph_create(const isph_params *params, int64_t n_cap, int64_t n_solid_cap,
                isph_ctx **out);
int isph_destroy(isph_ctx *ctx);
const char *isph_last_error(const isph_ctx *ctx);
int isph_set_params(isph_ctx *ctx, const isph_params *params);

/* ---- state transfer (caller-owned host SoA; copied) ----
 * isph_upload bins the particles into cells (half-open cell bounds as in
 * internal.f90:82), walls before fluid in each cell, input order kept
 * (boundary.f90:159-185 then internal.f90:80-91).  flags: ISPH_F_* bits. */
int isph_upload(isph_ctx *ctx, int64_t n, const double *x, const double *y,
                const double *u, const double *v, const double *p,
                const double *m, const double *rho, const double *rho0,
                const double *vol0, const double *c, const int32_t *tid,
                const uint8_t *flags);
/* Static porous-solid particles (the porlist of boundary.f90:523-593). */
int isph_upload_solids(isph_ctx *ctx, int64_t ns, const double *x,
                       const double *y, const double *m, const double *rho);
/* Download in the current cell-sorted order; any pointer may be NULL. */
int isph_download(isph_ctx *ctx, int64_t *n, double *x, double *y, double *u,
                  double *v, double *p, double *m, double *rho, double *rho0,
                  double *vol0, double *c, int32_t *tid, uint8_t *flags);
int isph_get_scalars(isph_ctx *ctx, isph_scalars *out);
int isph_set_scalars(isph_ctx *ctx, const isph_scalars *in);

/* ---- one entry per reference subroutine (main.f90:210-260) ---- */
int isph_projection(isph_ctx *ctx);              /* isph.f90:17-47          */
int isph_cellshift(isph_ctx *ctx);               /* integrator.f90:251-392  */
int isph_resetid(isph_ctx *ctx);                 /* isph.f90:88-113         */
int isph_neighbour(isph_ctx *ctx);               /* search.f90:13-399       */
int isph_effpor(isph_ctx *ctx);                  /* porous.f90:32-224       */
int isph_eddyvis(isph_ctx *ctx);                 /* turbulence.f90:12-110   */
int isph_int_vel(isph_ctx *ctx);                 /* integrator.f90:459-703  */
int isph_ppesolve(isph_ctx *ctx, isph_solve_stats *stats); /* isph.f90:115-488 */
int isph_comp_vel(isph_ctx *ctx);                /* integrator.f90:705-843  */
int isph_comp_pos(isph_ctx *ctx);                /* integrator.f90:845-883  */
int isph_cellshiftalt(isph_ctx *ctx);            /* integrator.f90:33-249   */
int isph_freesurf(isph_ctx *ctx);                /* isph.f90:49-86          */
int isph_opt2_shift(isph_ctx *ctx);              /* part_shift.f90:12-191   */
int isph_densityupdate(isph_ctx *ctx);           /* scalar.f90:399-426      */
int isph_timestep(isph_ctx *ctx);                /* integrator.f90:885-953  */
int isph_scalart_anisotropic(isph_ctx *ctx);     /* scalar.f90:242-362      */
int isph_scalarupdate(isph_ctx *ctx, double stime); /* scalar.f90:364-397   */

/* ---- inflow/outflow buffer zones (internal.f90:447-586, main.f90:217-235) ----
 * isph_set_buffer sets zone 1 or 2: the template particles (x, y, u, v, p,
 * in the reference's loop order buffer(k,i), k fastest) whose cells are the
 * zone's entry cells; the values an inserted particle takes (mass, density
 * -- also its reference density --, volume, concentration c_ref; the
 * reference's insert_buffer1 uses fmass, rho, fmass/rho, -0.5 and
 * insert_buffer2 fmass*relden, rho*relden, fmass/rho, +0.5); the removal
 * coordinate rv (zone 1 drops buffer particles with x >= rv, the reference's
 * rv_buf_r; zone 2 those with x <= rv, rv_buf_l); and the insertion clock:
 * isph_step updates the zone after the scalar sub-steps of the step in which
 * t + time_shift crosses iter * interval (ins_1, ins_2), iter = 1, 2, ...
 * interval 0 leaves the updates to isph_buffer_update.
 * isph_buffer_update runs remove_buffer then insert_buffer of the zone: in
 * each entry cell the kept particles keep their order and the zone's
 * template particles of the cell follow in template order.  Porosity moves
 * with its particle (1 for inserted ones); the capacity must hold the
 * current count plus the template. */
int isph_set_buffer(isph_ctx *ctx, int32_t zone, int64_t nt, const double *x,
                    const double *y, const double *u, const double *v,
                    const double *p, double mass, double rho, double vol0,
                    double c, double rv, double interval, double time_shift);
int isph_buffer_update(isph_ctx *ctx, int32_t zone);

/* ---- weakly compressible variant (wcsph.f90:13-154) ----
 * Dead and not compiling in the reference; restated from its equations.
 * isph_set_wcsph sets the sound speed co, the Tait exponent t_gam (7 in the
 * reference), the delta-SPH coefficient delt (0.1) and the still-water level
 * swl of the hydrostatic density (brrealy*distfac*(2*bl-1) + wc); with
 * in_step != 0, isph_step replaces the PPE assembly and solve by
 * calmasscons, updateden, dentopress and skips densityupdate (the density
 * is a state variable).  One entry per reference subroutine:
 *   presstoden   density from pressure, rho ((p gam/(rho co^2)) + 1)^(1/gam)
 *   dentopress   pressure from density (fluid), clipped at 0
 *   calmasscons  continuity + delta-SPH term into tden (fluid); the
 *                reference's y-component sum, overwritten at every
 *                neighbour by a variable reuse, has its own accumulator
 *   updateden    density = tden (fluid) */
int isph_set_wcsph(isph_ctx *ctx, double co, double t_gam, double delt, double swl,
                   int32_t in_step);
int isph_presstoden(isph_ctx *ctx);
int isph_dentopress(isph_ctx *ctx);
int isph_calmasscons(isph_ctx *ctx);
int isph_updateden(isph_ctx *ctx);

/* ---- piston wavemaker (no counterpart in the reference) ----
 * The paddle is the wall particles flagged ISPH_F_PADDLE.  It moves along x
 * by X(t) from its rest position, with velocity U(t):
 *   tau = t - t0;  a = stroke / 2;  w = 2 pi / period
 *   r   = ramp > 0 and tau < ramp ? tau / ramp : 1   (0 for tau <= 0)
 *   r'  = ramp > 0 and 0 < tau < ramp ? 1 / ramp : 0
 *   X   = (a r) sin(w tau)
 *   U   = a ((r w) cos(w tau) + r' sin(w tau))
 * isph_wavemaker_motion evaluates X and U at time t (host arithmetic, in
 * this order of operations).
 * isph_set_wavemaker records each paddle particle's rest position as its
 * current x minus X at the context's current time, and turns the motion on
 * (stroke 0 holds the paddle at rest).  isph_step then runs isph_wavemaker
 * after projection in every step.
 * isph_wavemaker moves the paddle to X at the context's current time:
 * x = rest + X, u = U, v = 0, and the intermediate velocity (which the PPE
 * source term reads for wall neighbours) us = U, vs = 0.  The rebinning classifies paddle particles by
 * position as it does fluid (the cellshiftalt cut-off aside), so they move
 * between cells; everything else treats them as the walls they are. */
int isph_wavemaker_motion(double stroke, double period, double ramp, double t0, double t,
                          double *x, double *u);
int isph_set_wavemaker(isph_ctx *ctx, double stroke, double period, double ramp, double t0);
int isph_wavemaker(isph_ctx *ctx);

/* The two halves of ppesolve, for phase-by-phase parity:
 * assembly (isph.f90:123-456) and solve + pressure scatter (isph.f90:462-486). */
int isph_ppe_assemble(isph_ctx *ctx);
int isph_ppe_solve(isph_ctx *ctx, isph_solve_stats *stats);

/* ---- fused entry points ----
 * isph_init_state: first binning + resetid + neighbour + effpor + eddyvis, the
 * state main.f90 holds when its production loop starts.
 * isph_step: nsteps production-loop iterations (main.f90:199-261) in the
 * reference's order; the stats of the last solve go to *stats (may be NULL). */
int isph_init_state(isph_ctx *ctx);
int isph_step(isph_ctx *ctx, int32_t nsteps, isph_solve_stats *stats);

/* ---- parity / debug dumps (device -> host copies of derived arrays) ----
 * Arrays are in the current sorted order.  Any pointer may be NULL. */
int isph_dump_cells(isph_ctx *ctx, int32_t *cell_of_particle,
                    int32_t *cell_start /* cellx*celly+1 */);
int isph_dump_neighbours(isph_ctx *ctx, int32_t *count /* n */,
                         int32_t *list /* n*ISPH_NBR_CAP, row-major */);
int isph_dump_props(isph_ctx *ctx, double *porosity, double *resist,
                    double *gradvx, double *coff /* 4n */, double *nut,
                    double *tke, double *varts, double *cdiff, double *vel,
                    double *us, double *vs, double *xs, double *ys);
/* PPE rows in canonical order: row_len[n], col/val [n*(ISPH_NBR_CAP+2)] row-major, rhs[n]. */
int isph_dump_ppe(isph_ctx *ctx, int32_t *row_len, int32_t *col, double *val,
                  double *rhs);
/* Upload an external PPE right-hand side / initial guess and solve only
 * (for solver tests on the assembled matrix).  The initial guess is the
 * FP64 form's (like the reference's fsol); the precision form starts every
 * solve from x = 0. */
int isph_ppe_set_rhs(isph_ctx *ctx, const double *rhs, const double *x0);
int isph_ppe_get_solution(isph_ctx *ctx, double *x);
/* Load a whole PPE system (n rows in the canonical row-major layout of
 * isph_dump_ppe, n <= capacity) in place of the assembled one, for solver
 * tests on systems assembled elsewhere.  Columns must lie in [0, n). */
int isph_load_system(isph_ctx *ctx, int32_t n, const int32_t *row_len,
                     const int32_t *col, const double *val, const double *rhs,
                     const double *x0);

/* ---- batched form: many independent scenarios stepped together ----
 * No counterpart in the reference (a single-scenario program); this is the
 * batched configuration's entry.  The contexts keep their own state; while
 * they belong to a batch they share one stream, and the PPE solves of the
 * precision-form contexts run as one set of launches over all of them.  A
 * scenario's results are the same as when it is stepped alone.  Destroy the
 * batch before any of its contexts.  isph_batch_step returns the worst
 * non-fatal solve code (ISPH_ENOCONV) over the scenarios and steps;
 * isph_batch_stats gives scenario s's last solve. */
typedef struct isph_batch isph_batch;
int isph_batch_create(int32_t n_scen, isph_ctx *const *ctxs, isph_batch **out);
int isph_batch_step(isph_batch *b, int32_t nsteps);
int isph_batch_stats(isph_batch *b, int32_t s, isph_solve_stats *out);
int isph_batch_destroy(isph_batch *b);
/* Kernel launches of the last step of a batch whose particle phases run
 * batched (one launch per phase over every scenario: every context is an
 * ISPH one with no insertion clock on a buffer zone), 0 otherwise.  While
 * a context belongs to a batch its solve form and the weakly compressible
 * switch are fixed (isph_set_params / isph_set_wcsph return ISPH_ESTATE). */
int64_t isph_batch_launches(isph_batch *b);
/* Live particle count of scenario s after the last batched step. */
int isph_batch_count(isph_batch *b, int32_t s, int64_t *n);
/* on != 0: the batched phase sequences before and after the solve run as
 * HIP graphs, captured at their first step (and again when a scenario's
 * launch shape changes) and replayed after; results are the same.  Only
 * for batches whose phases run batched (ISPH_ESTATE otherwise). */
int isph_batch_set_graph(isph_batch *b, int32_t on);
int isph_batch_autotune(isph_batch *b, int32_t trials, int32_t *shape, double *us);

/* ---- kernel timing (HIP events on the context's stream) ----
 * When enabled, every launch of the solver's timed kernels is bracketed by
 * a pair of HIP events; isph_prof_read returns the summed event time and the
 * launch count per kernel (launches after the solve has finished are left
 * out).  kernel: "spmv_p" (t = A p), "spmv_s" (t = A s) of the FP64 form;
 * "spmv_p32", "spmv_s32" (t1 = A ph, t2 = A sh), "mg_pre0", "mg_post0" (the
 * V-cycle's level-0 passes), "mg_restrict0" (the restriction to level 1),
 * "mg_down", "mg_up" (the multi-workgroup cell levels, every level),
 * "mg_coarse" (its single-workgroup coarse levels) and "pfinalize" (the
 * iteration's two dot-product finalizes) of the precision form. */
int isph_prof_enable(isph_ctx *ctx, int32_t on);
int isph_prof_read(isph_ctx *ctx, const char *kernel, double *total_ms,
                   int64_t *launches);

/* ---- autotuner (no counterpart in the reference) ----
 * Times the particle kernels below on the context's current state in each
 * of their launch shapes (trials launches each, HIP events on its stream)
 * and launches each in its fastest shape from then on.  Shapes: 0, 1, 2 =
 * one thread per particle in blocks of 64, 128, 256 threads; for the
 * neighbour search also 3, 4 = one workgroup per tile of 4 (8) cells of a
 * cell column on 128 (256) threads, the tile's neighbourhood staged in LDS.
 * Results do not depend on the shape.  Call it at a step boundary (after
 * isph_init_state or a step): the timed launches leave the state as it was
 * apart from values every step recomputes before reading them.  shape
 * (ISPH_TUNE_N, may be NULL) gets the chosen shapes; us (ISPH_TUNE_C *
 * ISPH_TUNE_N, may be NULL) the mean microseconds per launch of each shape,
 * -1 where a shape does not apply.  Kernel order: neighbour, effpor,
 * eddyvis, int_vel_varts, int_vel, ppe_assemble, comp_vel, opt2_shift,
 * scalart.  isph_batch_autotune does the same for the batched phases of a
 * batch (one launch over every scenario). */
#define ISPH_TUNE_N 9
#define ISPH_TUNE_C 5
int isph_autotune(isph_ctx *ctx, int32_t trials, int32_t *shape, double *us);
/* Launch tuned kernel `kernel` (the order above) in shape `shape` from now on
 * (ISPH_EARG for a shape that does not apply to it). */
int isph_set_shape(isph_ctx *ctx, int32_t kernel, int32_t shape);
int isph_batch_set_shape(isph_batch *b, int32_t kernel, int32_t shape);

/* Device memory a context of this capacity needs, in bytes (for admission). */
int64_t isph_footprint_bytes(int64_t n_cap, int64_t n_solid_cap,
                             int32_t cellx, int32_t celly, int32_t precision);

/* ---- allocation-free FP64 solve ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the FP64 BiCGStab of isph_ppe_solve in the
 * ISPH_PPE_FP64 form (same kernels, same reduction order, same restarts: the
 * iterates equal those of isph_load_system + isph_ppe_solve bit for bit) on
 * a system the caller holds in device memory, with no context.
 *   n rows (0 <= n <= ISPH_NOALLOC_MAX_ROWS), ELL slot-major with leading
 *   dimension ld >= n: entry k of row i is col[k*ld + i], val[k*ld + i] for
 *   k < row_len[i] <= ISPH_NBR_CAP + 2; rhs[n]; x[n] holds the initial guess
 *   on entry and the solution on return.  All five pointers are device
 *   memory.  tol: relative residual target; max_iter, max_outer: iteration
 *   and breakdown-restart caps (as isph_params).  stream: a hipStream_t, or
 *   NULL for the null stream; the call returns after the stream has drained.
 * The BiCGStab vectors, the block partials and the solver scalars live in a
 * static device array of fixed size (6 * ISPH_NOALLOC_MAX_ROWS doubles plus
 * ~64 KB, about 50 MB, held on every GPU the library runs on until the
 * process ends); the call makes no hipMalloc, hipFree or other allocation.
 * All calls on one GPU share that array, so calls that overlap in time on
 * one GPU (from two host threads, or on two streams) race and must not be
 * made.  Rows longer than ISPH_NBR_CAP + 2 or columns outside [0, n) are
 * rejected (ISPH_EARG) by a device check before the first product; n above
 * the maximum is rejected before any device call.  Errors are reported by
 * isph_last_error(NULL).  Return codes as isph_ppe_solve. */
#define ISPH_NOALLOC_MAX_ROWS (1 << 20)
int isph_ppe_solve_noalloc(int32_t n, int32_t ld, const int32_t *row_len,
                           const int32_t *col, const double *val, const double *rhs,
                           double *x, double tol, int32_t max_iter, int32_t max_outer,
                           void *stream, isph_solve_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* ISPH_ABI_H */
