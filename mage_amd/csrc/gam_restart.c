/*
 * gam_restart.c -- GAMERA restart dumps for the device step (libgamres).
 *
 * Host code over the HDF5 C library.  The file is the one writeH5Res
 * writes and readH5Restart reads (src/gamera/gioH5.F90:636-858): scalar
 * root attributes nOut, nRes, ts, t, ot, dt, dt0 and double datasets X, Y,
 * Z, Gas, magFlux, Bxyz, oGas, omagFlux, oBxyz, each dataset with "Units"
 * and "Description" string attributes (ioH5.F90:760-814, 1132-1184).  The
 * reference writes through the Fortran HDF5 interface, which stores an
 * array A(n1,...,nr) as a dataset of dimensions (nr,...,n1) with the same
 * bytes; the C calls below spell out that reversal.
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hdf5.h>
#include <hdf5_hl.h>

#include "gamres.h"

static __thread char g_err[512];

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

const char *gam_res_last_error(void) { return g_err; }

/* Fortran extents of one restart */
typedef struct {
    hsize_t ni, nj, nk;     /* cells incl. ghosts */
    hsize_t nip, njp, nkp;  /* active cells */
    int nG;
} Ext;

static int extents(const gam_dims *d, Ext *e)
{
    if (!d || d->Nip < 1 || d->Njp < 1 || d->Nkp < 1 || d->nG < 0)
        return fail(GAM_E_ARG, "gam_res: bad dimensions");
    e->nG = d->nG;
    e->nip = (hsize_t)d->Nip; e->njp = (hsize_t)d->Njp; e->nkp = (hsize_t)d->Nkp;
    e->ni = e->nip + 2 * (hsize_t)d->nG;
    e->nj = e->njp + 2 * (hsize_t)d->nG;
    e->nk = e->nkp + 2 * (hsize_t)d->nG;
    return GAM_OK;
}

static int put_attr(hid_t f, const char *name, hid_t type, const void *v)
{
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate2(f, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
    herr_t r = (a < 0) ? -1 : H5Awrite(a, type, v);
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return r < 0 ? fail(GAM_E_IO, "gam_res_write: attribute %s", name) : GAM_OK;
}

/* dataset of Fortran extents fdims[0..r-1] */
static int put_set(hid_t f, const char *name, int r, const hsize_t *fdims, const double *v)
{
    hsize_t cd[8];
    for (int q = 0; q < r; q++) cd[q] = fdims[r - 1 - q];
    if (H5LTmake_dataset_double(f, name, r, cd, v) < 0)
        return fail(GAM_E_IO, "gam_res_write: dataset %s", name);
    if (H5LTset_attribute_string(f, name, "Units", "") < 0 ||
        H5LTset_attribute_string(f, name, "Description", "") < 0)
        return fail(GAM_E_IO, "gam_res_write: attributes of %s", name);
    return GAM_OK;
}

int gam_res_write(const char *path, const gam_dims *dims, const gam_res_meta *m,
                  const double *xyz, const double *Gas, const double *magFlux,
                  const double *Bxyz, const double *oGas, const double *omagFlux,
                  const double *oBxyz)
{
    Ext e;
    int rc = extents(dims, &e);
    if (rc) return rc;
    if (!path || !m || !Gas || !magFlux || !oGas || !omagFlux)
        return fail(GAM_E_ARG, "gam_res_write: NULL argument");
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(GAM_E_IO, "gam_res_write: cannot create %s", path);
    /* writeH5Res order: nOut, nRes, ts, dt, dt0, X/Y/Z, State, oState */
    const double dt0 = (m->dt0 < 1.0e-12 * 10.0) ? 0.0 : m->dt0;   /* gioH5.F90:650-654 */
    const hsize_t cn[3] = {e.ni + 1, e.nj + 1, e.nk + 1};
    const hsize_t gas[5] = {e.ni, e.nj, e.nk, 5, 1};
    const hsize_t mf[4] = {e.ni + 1, e.nj + 1, e.nk + 1, 3};
    const hsize_t bx[4] = {e.ni, e.nj, e.nk, 3};
    const size_t ncorn = (size_t)cn[0] * cn[1] * cn[2];
    if (!rc) rc = put_attr(f, "nOut", H5T_NATIVE_INT, &m->nOut);
    if (!rc) rc = put_attr(f, "nRes", H5T_NATIVE_INT, &m->nRes);
    if (!rc) rc = put_attr(f, "ts", H5T_NATIVE_INT, &m->ts);
    if (!rc) rc = put_attr(f, "dt", H5T_NATIVE_DOUBLE, &m->dt);
    if (!rc) rc = put_attr(f, "dt0", H5T_NATIVE_DOUBLE, &dt0);
    if (!rc && xyz) rc = put_set(f, "X", 3, cn, xyz);
    if (!rc && xyz) rc = put_set(f, "Y", 3, cn, xyz + ncorn);
    if (!rc && xyz) rc = put_set(f, "Z", 3, cn, xyz + 2 * ncorn);
    if (!rc) rc = put_attr(f, "t", H5T_NATIVE_DOUBLE, &m->t);
    if (!rc) rc = put_set(f, "Gas", 5, gas, Gas);
    if (!rc) rc = put_set(f, "magFlux", 4, mf, magFlux);
    if (!rc && Bxyz) rc = put_set(f, "Bxyz", 4, bx, Bxyz);
    if (!rc) rc = put_attr(f, "ot", H5T_NATIVE_DOUBLE, &m->ot);
    if (!rc) rc = put_set(f, "oGas", 5, gas, oGas);
    if (!rc) rc = put_set(f, "omagFlux", 4, mf, omagFlux);
    if (!rc && oBxyz) rc = put_set(f, "oBxyz", 4, bx, oBxyz);
    if (H5Fclose(f) < 0 && !rc) rc = fail(GAM_E_IO, "gam_res_write: closing %s", path);
    return rc;
}

static int get_iattr(hid_t f, const char *name, int *v)
{
    if (H5LTget_attribute_int(f, "/", name, v) < 0)
        return fail(GAM_E_IO, "gam_res_read: attribute %s missing", name);
    return GAM_OK;
}

static int get_dattr(hid_t f, const char *name, double *v)
{
    if (H5LTget_attribute_double(f, "/", name, v) < 0)
        return fail(GAM_E_IO, "gam_res_read: attribute %s missing", name);
    return GAM_OK;
}

/* Read dataset `name` of Fortran extents full[] (with ghosts) or act[]
 * (without, older files) into out, laid out with ghosts; gofs = offset of
 * the first active element along the three spatial axes. */
static int get_set(hid_t f, const char *name, int r, const hsize_t *full, const hsize_t *act,
                   int nG, double *out)
{
    int rank = 0;
    if (H5LTget_dataset_ndims(f, name, &rank) < 0 || rank != r)
        return fail(GAM_E_IO, "gam_res_read: dataset %s missing or of rank %d", name, rank);
    hsize_t cd[8];
    H5T_class_t cls;
    size_t sz;
    if (H5LTget_dataset_info(f, name, cd, &cls, &sz) < 0 || cls != H5T_FLOAT)
        return fail(GAM_E_IO, "gam_res_read: dataset %s is not floating point", name);
    int isFull = 1, isAct = 1;
    for (int q = 0; q < r; q++) {
        if (cd[q] != full[r - 1 - q]) isFull = 0;
        if (cd[q] != act[r - 1 - q]) isAct = 0;
    }
    if (isFull) {
        if (H5LTread_dataset_double(f, name, out) < 0)
            return fail(GAM_E_IO, "gam_res_read: reading %s", name);
        return GAM_OK;
    }
    if (!isAct)
        return fail(GAM_E_ARG, "gam_res_read: %s was written for a grid of another size", name);
    /* restart without ghosts (gioH5.F90:771-773): fill the active block */
    size_t na = 1, nf = 1;
    for (int q = 0; q < r; q++) { na *= act[q]; nf *= full[q]; }
    double *tmp = (double *)malloc(na * sizeof(double));
    if (!tmp) return fail(GAM_E_NOMEM, "gam_res_read: out of host memory");
    if (H5LTread_dataset_double(f, name, tmp) < 0) {
        free(tmp);
        return fail(GAM_E_IO, "gam_res_read: reading %s", name);
    }
    size_t ncomp = 1;
    for (int q = 3; q < r; q++) ncomp *= act[q];
    for (size_t v = 0; v < ncomp; v++)
        for (hsize_t k = 0; k < act[2]; k++)
            for (hsize_t j = 0; j < act[1]; j++)
                for (hsize_t i = 0; i < act[0]; i++) {
                    size_t s = i + act[0] * (j + act[1] * (k + act[2] * v));
                    size_t d = (i + nG) + full[0] * ((j + nG) + full[1] * ((k + nG) + full[2] * v));
                    out[d] = tmp[s];
                }
    free(tmp);
    (void)nf;
    return GAM_OK;
}

int gam_res_read(const char *path, const gam_dims *dims, gam_res_meta *m, double *Gas,
                 double *magFlux, double *oGas, double *omagFlux)
{
    Ext e;
    int rc = extents(dims, &e);
    if (rc) return rc;
    if (!path || !m || !Gas || !magFlux || !oGas || !omagFlux)
        return fail(GAM_E_ARG, "gam_res_read: NULL argument");
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(GAM_E_IO, "gam_res_read: unable to open restart file %s", path);
    if (H5LTfind_dataset(f, "oGas") <= 0) {
        H5Fclose(f);
        return fail(GAM_E_IO, "gam_res_read: restart file too old, does not include oState");
    }
    const hsize_t gasF[5] = {e.ni, e.nj, e.nk, 5, 1}, gasA[5] = {e.nip, e.njp, e.nkp, 5, 1};
    const hsize_t mfF[4] = {e.ni + 1, e.nj + 1, e.nk + 1, 3}, mfA[4] = {e.nip + 1, e.njp + 1, e.nkp + 1, 3};
    if (!rc) rc = get_iattr(f, "nOut", &m->nOut);
    if (!rc) rc = get_iattr(f, "nRes", &m->nRes);
    if (!rc) rc = get_iattr(f, "ts", &m->ts);
    if (!rc) rc = get_dattr(f, "t", &m->t);
    if (!rc) rc = get_dattr(f, "dt", &m->dt);
    if (!rc) rc = get_dattr(f, "ot", &m->ot);
    if (!rc) {
        /* dt0 is optional; below 10 TINY it reads as 0 (gioH5.F90:814-820) */
        if (H5Aexists(f, "dt0") > 0) rc = get_dattr(f, "dt0", &m->dt0);
        else m->dt0 = 0.0;
        if (!rc && m->dt0 < 1.0e-12 * 10.0) m->dt0 = 0.0;
    }
    if (!rc) rc = get_set(f, "Gas", 5, gasF, gasA, e.nG, Gas);
    if (!rc) rc = get_set(f, "magFlux", 4, mfF, mfA, e.nG, magFlux);
    if (!rc) rc = get_set(f, "oGas", 5, gasF, gasA, e.nG, oGas);
    if (!rc) rc = get_set(f, "omagFlux", 4, mfF, mfA, e.nG, omagFlux);
    H5Fclose(f);
    return rc;
}

/* ------------------------------------------------------ device contexts */
static int dims_of(gam_ctx *ctx, gam_dims *d);

int gam_write_restart(gam_ctx *ctx, const char *path, const double *xyz, int nOut, int nRes)
{
    gam_dims d;
    int rc = dims_of(ctx, &d);
    if (rc) return rc;
    const size_t nc = (size_t)(d.Nip + 2 * d.nG) * (d.Njp + 2 * d.nG) * (d.Nkp + 2 * d.nG);
    const size_t nf = (size_t)(d.Nip + 2 * d.nG + 1) * (d.Njp + 2 * d.nG + 1) * (d.Nkp + 2 * d.nG + 1);
    double *buf = (double *)malloc(sizeof(double) * (2 * (5 * nc + 3 * nf + 3 * nc)));
    if (!buf) return fail(GAM_E_NOMEM, "gam_write_restart: out of host memory");
    double *G = buf, *M = G + 5 * nc, *B = M + 3 * nf;
    double *oG = B + 3 * nc, *oM = oG + 5 * nc, *oB = oM + 3 * nf;
    gam_res_meta m;
    m.nOut = nOut;
    m.nRes = nRes;
    if ((rc = gam_get_time(ctx, &m.t, &m.dt, &m.ts)) ||
        (rc = gam_get_dt0(ctx, &m.dt0)) ||
        (rc = gam_get_state(ctx, GAM_STATE, G, M, B, NULL, &m.t)) ||
        (rc = gam_get_state(ctx, GAM_OSTATE, oG, oM, oB, NULL, &m.ot))) {
        fail(rc, "gam_write_restart: %s", gam_last_error());
    } else {
        rc = gam_res_write(path, &d, &m, xyz, G, M, B, oG, oM, oB);
    }
    free(buf);
    return rc;
}

int gam_read_restart(gam_ctx *ctx, const char *path, int *nOut, int *nRes, double *dt)
{
    gam_dims d;
    int rc = dims_of(ctx, &d);
    if (rc) return rc;
    const size_t nc = (size_t)(d.Nip + 2 * d.nG) * (d.Njp + 2 * d.nG) * (d.Nkp + 2 * d.nG);
    const size_t nf = (size_t)(d.Nip + 2 * d.nG + 1) * (d.Njp + 2 * d.nG + 1) * (d.Nkp + 2 * d.nG + 1);
    double *buf = (double *)calloc(2 * (5 * nc + 3 * nf), sizeof(double));
    if (!buf) return fail(GAM_E_NOMEM, "gam_read_restart: out of host memory");
    double *G = buf, *M = G + 5 * nc, *oG = M + 3 * nf, *oM = oG + 5 * nc;
    gam_res_meta m;
    rc = gam_res_read(path, &d, &m, G, M, oG, oM);
    if (!rc) {
        double dtNew = 0.0;
        /* readH5Restart: State, oState, ts, t, dt, dt0; then the restart
         * branch of the set-up: bFlux2Fld of both states, Model%dt = CalcDT
         * (init.F90:205-246) */
        if ((rc = gam_put_state(ctx, GAM_STATE, G, M, NULL, NULL, m.t)) ||
            (rc = gam_put_state(ctx, GAM_OSTATE, oG, oM, NULL, NULL, m.ot)) ||
            (rc = gam_set_time(ctx, m.t, m.dt, m.ts)) ||
            (rc = gam_set_dt0(ctx, m.dt0)) ||
            (rc = gam_bflux2fld(ctx, GAM_STATE)) ||
            (rc = gam_bflux2fld(ctx, GAM_OSTATE)) ||
            (rc = gam_calc_dt(ctx, &dtNew)) ||
            (rc = gam_set_time(ctx, m.t, dtNew, m.ts))) {
            fail(rc, "gam_read_restart: %s", gam_last_error());
        } else {
            if (nOut) *nOut = m.nOut;
            if (nRes) *nRes = m.nRes + 1;
            if (dt) *dt = dtNew;
        }
    }
    free(buf);
    return rc;
}

static int dims_of(gam_ctx *ctx, gam_dims *d)
{
    if (!ctx) return fail(GAM_E_ARG, "gam_res: NULL context");
    if (gam_get_dims(ctx, d)) return fail(GAM_E_ARG, "gam_res: %s", gam_last_error());
    return GAM_OK;
}
