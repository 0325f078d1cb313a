/*
 * gamres.h -- GAMERA restart files for the MI355X step (libgamres).
 *
 * Writes and reads the reference's restart dump <runid>.gam.Res.<nRes>.h5
 * (writeH5Res / readH5Restart, src/gamera/gioH5.F90:636-858; file name
 * output.F90:152-172), so a run can move between the reference and this
 * build in either direction, and so that checkpoints of the device step
 * have a defined on-disk form.
 *
 * File layout (HDF5, written as the reference's Fortran HDF5 calls produce
 * it, i.e. dataset dimensions reversed with respect to the Fortran arrays):
 *   root attributes (scalars): nOut, nRes, ts (int); t, ot, dt, dt0 (double)
 *   datasets (double): X, Y, Z        corners,      Fortran (Ni+1,Nj+1,Nk+1)
 *                      Gas, oGas      (Ni,Nj,Nk,5,1)
 *                      magFlux, omagFlux (Ni+1,Nj+1,Nk+1,3)
 *                      Bxyz, oBxyz    (Ni,Nj,Nk,3)
 *   every dataset carries the string attributes "Units" and "Description".
 * Ni = Nip + 2 nG etc.; arrays include the ghost layers.  The reader also
 * accepts the older layout without ghosts (gioH5.F90:767-776).
 *
 * Return codes are those of gamhip.h, plus GAM_E_IO for file errors; the
 * message of the last failure is gam_res_last_error().
 */
#ifndef GAMRES_H
#define GAMRES_H

#include "gamhip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GAM_E_IO -7

/* Model_T/State_T scalars of a restart */
typedef struct {
    int nOut, nRes, ts;
    double t, ot, dt, dt0;
} gam_res_meta;

const char *gam_res_last_error(void);

/* Host arrays <-> file (Fortran layouts as in gamhip.h).  Any of xyz and
 * the Bxyz arrays may be NULL on write (then they are not written); on
 * read, Gas/magFlux/oGas/omagFlux are filled (the reference reads no other
 * arrays: B is recomputed from the fluxes). */
int gam_res_write(const char *path, const gam_dims *dims, const gam_res_meta *meta,
                  const double *xyz, const double *Gas, const double *magFlux,
                  const double *Bxyz, const double *oGas, const double *omagFlux,
                  const double *oBxyz);
int gam_res_read(const char *path, const gam_dims *dims, gam_res_meta *meta,
                 double *Gas, double *magFlux, double *oGas, double *omagFlux);

/* WriteRestart of a device context (resOutput, output.F90:147-173
 * -> writeH5Res): State, oState, t, dt, ts and dt0 of the context; xyz =
 * the grid corners (Ni+1,Nj+1,Nk+1,3) or NULL. */
int gam_write_restart(gam_ctx *ctx, const char *path, const double *xyz, int nOut, int nRes);

/* Restart of a device context: readH5Restart followed by the restart branch
 * of the set-up tail (init.F90:190-244): State and oState (Gas, magFlux,
 * times), ts, t, dt0 from the file, B of both states from their fluxes, no
 * boundary pass, then Model%dt = CalcDT.  *nRes receives the file's nRes+1
 * as in the reference.  The grid (and background) must already be set. */
int gam_read_restart(gam_ctx *ctx, const char *path, int *nOut, int *nRes, double *dt);

#ifdef __cplusplus
}
#endif
#endif
