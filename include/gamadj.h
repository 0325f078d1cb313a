/*
 * gamadj.h -- discrete adjoint of the GAMERA step on the MI355X (libgamhip).
 *
 * No counterpart in the reference (it has no adjoint).  The adjoint is of
 * the step this library runs (stepGamera, gamapp.F90:73-113), built by
 * hand kernel by kernel in reverse order: boundary copies, the Reynolds
 * update with Flux2Deltas (mhdgroup.F90:79-121, stress.F90:152-194), the
 * interface fluxes (reconstruction with the PDM limiter's branch taken in
 * the forward run, GasKinFlux; stress.F90:204-511, recon.F90:125-301) and
 * the AB2 predictor (mhdgroup.F90:444-561).
 *
 * Scope: every step option: hydrodynamic and MHD runs (Maxwell or Boris
 * update, with or without the background field) with periodic,
 * zero-gradient, sphere and inflow (wind) boundaries.  With MHD the adjoint
 * runs through the edge EMFs, constrained transport, the magnetic kinetic
 * fluxes with HOGS (and the Boris mag-HOGS and nuke terms), the background
 * cross terms and CellMaxwell / CellBoris.  The inflow boundary's adjoint
 * (wind.F90:226-335) covers the ghost-gas blend with the fast-speed
 * outflow fix, the blended ghost face fluxes and the blend weight's
 * dependence on the wind direction, and yields the gradient with respect
 * to the inflow time series (gam_adj_wind_grad).
 *
 * Objective: J = sum over cells of w(c) * P(c) of the State after nsteps
 * (P the CellC2P pressure), w a weight array (Ni,Nj,Nk) given by the
 * caller (e.g. cell volume times a region mask).
 *
 * dt: by default each step keeps the dt of the recorded forward run (the
 * CFL minimum is not differentiated) and gam_adj_replay runs the forward
 * with those dt values.  gam_adj_set_dt_mode turns on the dt
 * sensitivities: dJ/d(dt_n) of every step (gam_adj_dt_grad), and, in mode
 * 2, the gradient of the run in which each dt after the first is CalcDT's
 * minimum of the state before it (step.F90:58-101), as in the reference's
 * step loop; gam_adj_replay_dt replays with a given or a free dt sequence.
 *
 * Memory: checkpoints of State and oState every ckpt_every steps, and the
 * states of one checkpoint interval, recomputed forward before its
 * adjoint steps run backward, with the interval's forward intermediates
 * (half state, face velocities, face fluxes) as far as HBM allows.  Where
 * it allows a second set of interval tapes, the next interval is
 * recomputed on a second stream (default priority) while the current one
 * runs backward.  gam_adj_forward leaves the last interval on the tapes.
 */
#ifndef GAMADJ_H
#define GAMADJ_H

#include "gamhip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gam_adj gam_adj;

/* An adjoint run of nsteps steps of ctx with checkpoints every ckpt_every
 * (1 <= ckpt_every <= nsteps). */
int gam_adj_create(gam_adj **out, gam_ctx *ctx, int nsteps, int ckpt_every);
/* Same, with the tape budget set by the caller: at most max_taped steps of
 * an interval keep their forward intermediates on the tape (max_taped < 0:
 * as many as HBM allows, as gam_adj_create; 0: every adjoint step
 * recomputes its predictor, face velocities and fluxes), and overlap = 0
 * keeps a single tape pair (the interval recompute then runs on the
 * context's stream before the interval's adjoint steps).  Gradients do not
 * depend on either choice; large grids take these paths when HBM is short. */
int gam_adj_create_budget(gam_adj **out, gam_ctx *ctx, int nsteps, int ckpt_every,
                          int max_taped, int overlap);
/* Steps of an interval whose intermediates are taped, and whether the
 * reverse sweep overlaps the interval recompute (second tape pair). */
int gam_adj_tape_info(gam_adj *a, int *taped_steps, int *overlap);
int gam_adj_destroy(gam_adj *a);
/* Objective weights w(Ni,Nj,Nk) (Fortran layout, ghosts included). */
int gam_adj_set_weights(gam_adj *a, const double *w);
/* nsteps x stepGamera from the context's current State, oState and clock,
 * recording checkpoints and every step's clock; *J = objective at the
 * end.  The context is left at the final state. */
int gam_adj_forward(gam_adj *a, double *J);
/* Reverse sweep of the recorded run: gGas = dJ/d(State Gas) and gGasO =
 * dJ/d(oState Gas) at the start, (Ni,Nj,Nk,5); with MHD gMF and gMFO =
 * dJ/d(magFlux) of State and oState, (Ni+1,Nj+1,Nk+1,3).  Any output but
 * gGas may be NULL.  Gradients are with respect to the active cells and
 * the face fluxes: each start state's ghost cells follow from its interior
 * through the boundary conditions and its cell B from its face fluxes
 * (bFlux2Fld), so those adjoints are folded in (the ghost entries are 0).
 * The start is taken in gam_adj_replay's order (active cell B, the
 * boundary pass, every cell's B), the form stepGamera leaves State in
 * after each step.  The reference's set-up (and gam_init_run) forms ghost
 * B by the boundary rules instead (init.F90:200-244); gGas does not depend
 * on the difference, gMF does: for gMF to be the gradient of the run,
 * start it from that form (gam_init_run, then gam_bflux2fld of State and
 * of oState).  The context is left at the final state of the forward run. */
int gam_adj_backward(gam_adj *a, double *gGas, double *gGasO, double *gMF, double *gMFO);
/* Forward from the recorded start with the given State Gas (and oState
 * Gas, State and oState magFlux; NULL keeps the recorded array), ghost
 * cells refilled by the boundary conditions (the oState's at oState%time,
 * which matters for the time-dependent inflow) and cell B recomputed from
 * the face fluxes, each step's dt held at the recorded value; *J = objective
 * at the end.  The context is then put back at the final state of the
 * recorded run. */
int gam_adj_replay(gam_adj *a, const double *Gas, const double *oGas, const double *MF,
                   const double *oMF, double *J);
/* After gam_adj_backward with an inflow BC: gQ(n,5) = dJ/dQ and gB(n,3) =
 * dJ/dB of the series given to gam_set_wind (column-major, the same
 * layout; the Bx column is 0: Bx = Bx0 + ByC*By + BzC*Bz).  Each sample
 * the boundary reads at a time counts; the sample index itself is a step
 * function of time and is not differentiated.  The series must be the one
 * the forward run used.  GAM_E_STATE without an inflow BC. */
int gam_adj_wind_grad(gam_adj *a, double *gQ, double *gB);
/* Time at which the start oState's ghost cells follow its interior through
 * the boundary conditions, in the adjoint and in gam_adj_replay:
 * at_state_time = 0 (default) at oState%time; 1 at State%time, which is
 * what gam_init_run produces (oState = State after the boundary pass,
 * init.F90:230-244), so a forward that starts with gam_init_run is
 * differentiated exactly.  The two differ only for the time-dependent
 * inflow BC.  Takes effect at the next gam_adj_forward. */
int gam_adj_set_ostate_bc_time(gam_adj *a, int at_state_time);
/* dt sensitivities of the next gam_adj_backward (no counterpart in the
 * reference): mode 0 (default) none, the dt of every step held at the
 * recorded value; 1: also dJ/d(dt_n) for n = 0..nsteps-1, each with the
 * other steps' dt held (the partial derivatives through the update, the
 * constrained transport, the EMF's vDiff cap CFL*edge/dt and the AB2
 * predictor weight dt_n / (2 dt_(n-1))); 2: dt_(n+1) = CalcDT of the state
 * step n leaves before its boundary pass (the minimum of CellDT over
 * is-1..ie+1 x js-1..je+1 x ks-1..ke+1, gridutils.F90:49-106), so the
 * gradients are those of the run whose dt follows the state, and
 * dJ/d(dt_n) is the total derivative (dt_0 is the caller's input); 3: as
 * 2, and dt_0 is CalcDT of the start State as gam_adj_forward finds it
 * (after its boundary pass, as gam_init_run computes it), so the
 * gradients are those of a run that starts from gam_init_run.  A dt
 * the recorded run capped (tFin, 10*dt0) does not follow the state.  Modes
 * 2 and 3 take effect at the next gam_adj_forward (its steps record CalcDT's
 * argmin cell and values before the boundary pass, the state that exists
 * only inside the step); modes 0 and 1 at the next gam_adj_backward.  The
 * first call with mode > 0 allocates 2 planes of the grid. */
int gam_adj_set_dt_mode(gam_adj *a, int mode);
/* After gam_adj_backward with the dt sensitivities on: g[n] = dJ/d(dt_n),
 * n = 0..nsteps-1. */
int gam_adj_dt_grad(gam_adj *a, double *g);
/* dt of each step of the recorded run, dts[n], n = 0..nsteps-1. */
int gam_adj_step_dt(gam_adj *a, double *dts);
/* gam_adj_replay with the dt of step n = dts[n] (nsteps values, each >= 0;
 * a 0 takes CalcDT's dt: of the state the step before left, or, for the
 * first step, of the start after its boundary pass, as gam_init_run does),
 * or, with dts NULL, the recorded dt for the first step and CalcDT's for
 * each later one (the forward the dt mode 2 gradients belong to; all zeros
 * is the forward of mode 3). */
int gam_adj_replay_dt(gam_adj *a, const double *Gas, const double *oGas, const double *MF,
                      const double *oMF, const double *dts, double *J);
/* Milliseconds of the last gam_adj_forward and gam_adj_backward calls
 * (HIP events on the context's stream). */
int gam_adj_times(gam_adj *a, double *fwd_ms, double *bwd_ms);

#ifdef __cplusplus
}
#endif
#endif
