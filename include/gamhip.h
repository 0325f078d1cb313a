/*
 * gamhip.h -- C ABI of the MI355X-native GAMERA MHD time step.
 *
 * The entry points replace the reference's Fortran operator interface for
 * the GAMERA advance step (rgorby/MAGE, src/gamera).  A Fortran host binds
 * them through ISO_C_BINDING (mage_amd/fortran/gamhip_f.F90); Python binds
 * them through ctypes (mage_amd/gamhip.py).  Signatures carry only plain
 * pointers, sizes and scalars.
 *
 * Conventions
 *  - Every call returns an int: GAM_OK (0) or a negative GAM_E* code; the
 *    message of the last failure on the calling thread is gam_last_error().
 *    Where the reference prints and stops (step.F90:145-151,
 *    init.F90:817-826, bcs.F90:196-206) these calls return an error code.
 *  - Host arrays are the reference's own Fortran arrays: column-major, i
 *    fastest, ghost lower bounds isg=1-nG.  The state, grid, background and
 *    adjoint-gradient arrays may also be device memory of the context's GPU
 *    in that same layout (the copies use hipMemcpyDefault).  Cell arrays are
 *    (Ni,Nj,Nk[,comp...]) with Ni=Nip+2nG, face/edge arrays are
 *    (Ni+1,Nj+1,Nk+1[,comp...]) exactly as allocated by gamutils.F90:169-201
 *    (allocState/allocGridVec) and init.F90:793-1072 (Corners2Grid).
 *  - One context owns one HIP stream; calls are asynchronous on it except
 *    the ones that return host data (get_*, calc_dt, step with non-NULL
 *    dt_next, run, sync).  Contexts are not thread-safe; one host thread per
 *    context, as in the reference (single host thread, OpenMP inside).
 *  - All arithmetic is FP64.
 */
#ifndef GAMHIP_H
#define GAMHIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define GAM_OK            0
#define GAM_E_ARG        -1   /* invalid argument / size */
#define GAM_E_HIP        -2   /* HIP runtime failure */
#define GAM_E_STATE      -3   /* call made in the wrong order */
#define GAM_E_DT_BAD     -4   /* CalcDT "bad" (step.F90:103-133) */
#define GAM_E_DT_DISASTER -5  /* CalcDT "disaster" (step.F90:97-101,135-139) */
#define GAM_E_NOMEM      -6

/* Boundary kinds (reference BC classes, bcs.F90 / wind.F90) */
#define GAM_BC_PERIODIC  0    /* periodic{Inner,Outer}{I,J,K}BC_T  bcs.F90:279-450 */
#define GAM_BC_ZEROGRAD  1    /* zeroGradient*BC_T                 bcs.F90:799-930 */
#define GAM_BC_SPHIN     2    /* sphereInBC_T  (inner i)           bcs.F90:934-957 */
#define GAM_BC_SPHOUT    3    /* sphereOutBC_T (outer i)           bcs.F90:959-982 */
#define GAM_BC_WIND      4    /* WindBC_T      (outer i)           wind.F90:206-374 */

/* Reconstruction (sim/rmeth, init.F90:442-454) */
#define GAM_RECON_8CENT  0
#define GAM_RECON_7UP    1

/* State selectors */
#define GAM_STATE        0    /* gamApp_T%State  */
#define GAM_OSTATE       1    /* gamApp_T%oState */
#define GAM_HFSTATE      2    /* gamSolver_T%StateHf (read-only) */

/* Grid_T sizes: Nip/Njp/Nkp and nG (gamtypes.F90:168-175, init.F90:641-656).
 * Nip must be >= 16 and a multiple of 8 (init.F90:811-826). */
typedef struct {
    int Nip, Njp, Nkp;
    int nG;
} gam_dims;

/* Model_T knobs on the hot path (gamtypes.F90:77-151; defaults
 * init.F90:312-484).  bc[] = external BC kind per face in the order
 * INI,OUTI,INJ,OUTJ,INK,OUTK (gdefs.F90:41-43). */
typedef struct {
    double gamma, CFL, Vd0, Ca, cHogH, cHogM, pdmb, dFloor, pFloor;
    double tFin, dt0, limDT0;
    int doMHD, doBoris, doBackground, doLFMLim, rmeth;
    int bc[6];
} gam_model;

/* Grid_T metric arrays (host, Fortran layout; see gamtypes.F90:194-230) */
typedef struct {
    double *xyzcc;   /* (Ni,Nj,Nk,3)              */
    double *xfc;     /* (Ni+1,Nj+1,Nk+1,3,3)      */
    double *di, *dj, *dk, *volume;   /* (Ni,Nj,Nk) */
    double *Tf;      /* (Ni+1,Nj+1,Nk+1,9,3)      */
    double *Te;      /* (Ni+1,Nj+1,Nk+1,9,3)      */
    double *Teb;     /* (Ni+1,Nj+1,Nk+1,4,3)      */
    double *face;    /* (Ni+1,Nj+1,Nk+1,3)        */
    double *edge;    /* (Ni+1,Nj+1,Nk+1,3)        */
} gam_grid_arrays;

/* Grid_T background-field arrays (background.F90:43-48) */
typedef struct {
    double *fcB0;    /* (Ni,Nj,Nk,3,3) */
    double *edgB0;   /* (Ni,Nj,Nk,2,3) */
    double *B0;      /* (Ni,Nj,Nk,3)   */
    double *bFlux0;  /* (Ni,Nj,Nk,3)   */
    double *dpB0;    /* (Ni,Nj,Nk,3)   */
} gam_background_arrays;

typedef struct gam_ctx gam_ctx;

/* Message of the last failing call on this thread. */
const char *gam_last_error(void);

/* Allocate device storage for one Grid/State/Solver set.
 * Replaces the allocations of Hatch/initSolver (init.F90:38-59, 486-508). */
int gam_create(gam_ctx **ctx, const gam_dims *dims, const gam_model *model);
int gam_destroy(gam_ctx *ctx);
/* The sizes the context was created with. */
int gam_get_dims(gam_ctx *ctx, gam_dims *dims);

/* Upload Grid_T metrics computed by the host (the output of Corners2Grid,
 * init.F90:793-1072). */
int gam_set_grid(gam_ctx *ctx, const gam_grid_arrays *g);
/* Compute the metrics on the device from corner coordinates
 * xyz(Ni+1,Nj+1,Nk+1,3) -- replaces Corners2Grid + ebGeom
 * (init.F90:793-1137, quadrature.F90:42-221). */
int gam_grid_from_corners(gam_ctx *ctx, const double *xyz);
/* Download the device metrics (any pointer may be NULL). */
int gam_get_grid(gam_ctx *ctx, const gam_grid_arrays *g);

/* Background field: upload (AddB0 output, background.F90:20-167) or
 * compute on the device for the dipole of msphutils.F90:534-551 with moment
 * (0,0,M0).  Requires the grid metrics. */
int gam_set_background(gam_ctx *ctx, const gam_background_arrays *b);
int gam_add_dipole_b0(gam_ctx *ctx, double M0);
int gam_get_background(gam_ctx *ctx, const gam_background_arrays *b);

/* Inflow time series for GAM_BC_WIND (wind.F90 readWind/InterpWind):
 * tW(n), Q(n,5) primitives (rho,vx,vy,vz,P), B(n,3); Bx = Bx0+ByC*By+BzC*Bz;
 * xyzW = position the lag is measured from (wind.F90:157-175). */
int gam_set_wind(gam_ctx *ctx, int n, const double *tW, const double *Q,
                 const double *B, double Bx0, double ByC, double BzC,
                 const double *xyzW);

/* State_T upload/download (gamtypes.F90:253-271).  Gas(Ni,Nj,Nk,5),
 * magFlux(Ni+1,Nj+1,Nk+1,3), Bxyz(Ni,Nj,Nk,3), Efld(Ni,Nj,Nk,3).
 * Any array pointer may be NULL to skip it. */
int gam_put_state(gam_ctx *ctx, int which, const double *Gas, const double *magFlux,
                  const double *Bxyz, const double *Efld, double time);
int gam_get_state(gam_ctx *ctx, int which, double *Gas, double *magFlux,
                  double *Bxyz, double *Efld, double *time);

/* Model_T time fields (t, dt, ts) */
int gam_set_time(gam_ctx *ctx, double t, double dt, int ts);
int gam_get_time(gam_ctx *ctx, double *t, double *dt, int *ts);
/* Model%dt0, the reference step of the CalcDT checks (step.F90:120-124,
 * 164-166); set from a restart file by readH5Restart (gioH5.F90:814-820). */
int gam_set_dt0(gam_ctx *ctx, double dt0);
int gam_get_dt0(gam_ctx *ctx, double *dt0);

/* Finish initialisation after the IC and background are in place: the
 * tail of CalcGridInfo (init.F90:200-244): bFlux2Fld, oState=State,
 * EnforceBCs, oState=State, dt=CalcDT, oState%time=t-dt. */
int gam_init_run(gam_ctx *ctx, double *dt);

/* Operator-level entry points of the GAMERA advance step */
int gam_advance_mhd(gam_ctx *ctx, double dt);      /* AdvanceMHD   mhdgroup.F90:16-76 */
int gam_calc_dt(gam_ctx *ctx, double *dt);         /* CalcDT       step.F90:58-168 */
int gam_enforce_bcs(gam_ctx *ctx);                 /* EnforceBCs   step.F90:17-55  */
int gam_bflux2fld(gam_ctx *ctx, int which);        /* bFlux2Fld    gridutils.F90:281-301 */

/* stepGamera (gamapp.F90:73-113).  dt_next may be NULL: then nothing is
 * copied back and the call does not synchronise. */
int gam_step(gam_ctx *ctx, double *dt_next);
/* nsteps x stepGamera without host synchronisation between steps; returns
 * the first CalcDT failure code seen, if any.  Pairs of steps are replayed
 * as one hipGraph of two steps, captured on first use for each State
 * buffer (a step is kernels only, so a replay launches exactly what
 * gam_step would: the results are bit-identical).  Inside the graph the
 * j and k interface fluxes and the EMFs that take their face velocities
 * are branches beside the i-direction ones, which the device may run
 * side by side (GAM_CONCURRENT=0 in the environment at gam_create keeps
 * the graph a single chain).  Changing the grid, background, inflow
 * series or launch plan drops the captured graphs; with profiling on,
 * steps are launched one by one. */
int gam_run(gam_ctx *ctx, int nsteps);
/* gamAdvanceModel (gamtypessub.F90:74-90): one stepGamera, then more while
 * Model%t < t0 + dt.  *nsteps (may be NULL) receives the step count.
 * Synchronises after every step (the loop tests Model%t on the host). */
int gam_advance_model(gam_ctx *ctx, double dt, int *nsteps);
int gam_sync(gam_ctx *ctx);

/* Total |div B| over active cells, background included
 * (DivB, gridutils.F90:717-770). */
int gam_divb(gam_ctx *ctx, double *totDivB);

/* gam_divb without a context and without any device allocation.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * (code under HIP graph capture, tight loops, harnesses that forbid
 * hipMalloc).  magFlux: (Ni+1,Nj+1,Nk+1,3) face fluxes and bFlux0: NULL or
 * the background's (Ni,Nj,Nk,3) face fluxes, both DEVICE memory of the
 * current GPU in the reference's Fortran layout; *totDivB (device memory)
 * receives the total |div B| over the active cells.  Kernels only, on
 * `stream` (a hipStream_t; NULL = the null stream), asynchronous.  The
 * per-block partial sums live in a fixed 16 KiB device array that each
 * GPU holds for the life of the process; every call overwrites all of it,
 * so nothing carries over between calls, but calls that overlap on one GPU
 * (two streams, two host threads) race on it.  The first call on a device
 * looks up that array's address (not allowed under graph capture); later
 * calls may be captured.  GAM_E_ARG for NULL arguments, nG outside 1..16,
 * or more than GAM_DIVB_NOALLOC_MAX_CELLS active cells. */
#define GAM_DIVB_NOALLOC_MAX_CELLS 2147483647L
int gam_divb_noalloc(const gam_dims *dims, const double *magFlux, const double *bFlux0,
                     void *stream, double *totDivB);

/* hipEvent-timed loop of nsteps x stepGamera on the context's stream (as
 * gam_run, graphs included).  ms = elapsed milliseconds between the events. */
int gam_time_steps(gam_ctx *ctx, int nsteps, double *ms);

/* Per-kernel-group timing: when enabled, every launch group of the step is
 * bracketed by hipEvents on the context's stream.  gam_kernel_times()
 * returns accumulated milliseconds and launch counts per group in the
 * order of gam_kernel_group_name(). */
#define GAM_KG_COUNT 8
int gam_set_profiling(gam_ctx *ctx, int on);
int gam_kernel_times(gam_ctx *ctx, double *ms, int *counts);
const char *gam_kernel_group_name(int group);

/* Solver intermediates for staged comparison: which = 0 gFlx (Ni,Nj,Nk,5,3),
 * 1 mFlx (Ni,Nj,Nk,3,3).  Values are those of the last AdvanceMHD. */
int gam_get_solver(gam_ctx *ctx, int which, double *out);

/* ---------------------------------------------------------------------
 * Launch plans.  No counterpart in the reference (its loop blocking is
 * fixed at compile time, vecLen in kdefs.F90).  The step's tunable
 * launches ("slots": predictor, the EMF of each edge direction, the
 * interface flux of each face direction, the update) each have a few
 * launch shapes that give the same results: the chunk tiling of the box's
 * points, value = row order (1 = j fastest, 2 = k fastest, 3 = slabs of 8
 * j-rows) + 8 x chunk extent (0: 256 points per workgroup, 1: 512,
 * 2: 128), and, for the j and k fluxes, the LDS-staged face tile of
 * 64 x 4 (4) or 64 x 8 (5) faces, or those tiles with their rows brought
 * to LDS by direct global-to-LDS loads (6: 64 x 4, 7: 64 x 8).
 * gam_autotune times every candidate of every slot on the context's own
 * grid (reps steps per candidate, hipEvents around that slot's launch),
 * keeps the fastest (the current shape unless another is more than 1 %
 * faster), keeps the resulting plan only if 4 x reps whole graph-replayed
 * steps also run more than 1 % faster with it than with the plan it
 * started from, and restores State, oState, time and dt as they were; the
 * solver intermediates (gam_get_solver) are then those of a tuning step.
 * ------------------------------------------------------------------- */
#define GAM_TUNE_SLOTS 8
#define GAM_TUNE_MAXREPS 32
int gam_autotune(gam_ctx *ctx, int reps);             /* 1 <= reps <= GAM_TUNE_MAXREPS */
int gam_get_plan(gam_ctx *ctx, int *plan);            /* GAM_TUNE_SLOTS values */
int gam_set_plan(gam_ctx *ctx, const int *plan);      /* GAM_E_ARG if a value is not a candidate */
const char *gam_tune_slot_name(int slot);              /* "" outside 0..GAM_TUNE_SLOTS-1 */
/* Candidate values of a slot (up to GAM_TUNE_MAXCAND); returns their number,
 * 0 for a bad slot. */
#define GAM_TUNE_MAXCAND 16
int gam_tune_candidates(int slot, int *values);

/* ---------------------------------------------------------------------
 * Ensembles: many independent runs stepped together.  The reference runs
 * one gamApp_T per process and has no batched form; an ensemble here is a
 * set of contexts (each one stepGamera instance, gamapp.F90:73-113, with
 * its own grid size, grid, background, inflow series, state and clock)
 * whose steps are issued as one launch per kernel covering every member,
 * and whose whole step is one captured hipGraph.  Every member's results
 * are those of stepping it on its own.
 * ------------------------------------------------------------------- */
typedef struct gam_ens gam_ens;

#define GAM_ENS_MAX 256
/* Batch n (1..GAM_ENS_MAX) contexts, each ready to step (grid, background,
 * wind as its BCs need, state, gam_init_run).  Members must share the
 * kernel variant: rmeth, doMHD, doBoris, doBackground, doLFMLim and the six
 * BC kinds (GAM_E_ARG otherwise).  The context handles stay valid and
 * owned by the caller; state transfers (gam_get_state, ...) on a member
 * show the ensemble's progress.  A member may also be stepped on its own
 * between ensemble runs only an even number of times (the recorded graphs
 * fix which of its two state buffers is current; GAM_E_STATE otherwise).
 * The recorded launches keep each member's grid, background, inflow series
 * and launch plans as they were at creation: create the ensemble again
 * after changing any of them (gam_ens_run returns GAM_E_STATE once a
 * member's gam_set_grid, gam_grid_from_corners, gam_set_background,
 * gam_add_dipole_b0, gam_set_wind, gam_set_plan or gam_autotune has run
 * after gam_ens_create). */
int gam_ens_create(gam_ens **out, int n, gam_ctx *const *members);
/* nsteps x stepGamera of every member; synchronises at the end and returns
 * the first member's CalcDT failure code, if any (gam_last_error names the
 * member). */
int gam_ens_run(gam_ens *e, int nsteps);
/* Same, bracketed by hipEvents on the ensemble's stream; ms = elapsed time. */
int gam_ens_time_steps(gam_ens *e, int nsteps, double *ms);
/* Kernel launches per step of every member (= graph nodes per step). */
int gam_ens_launches_per_step(gam_ens *e, int *n);
int gam_ens_destroy(gam_ens *e);

#ifdef __cplusplus
}
#endif
#endif
