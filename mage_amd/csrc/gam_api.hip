// gam_api.hip -- the C ABI (include/gamhip.h): context, layout conversion
// and the step driver.  Host code only; kernels live in gam_step.hip and
// gam_grid.hip.
#include "gamhip.h"
#include "gamadj.h"
#include "gam_kernels.h"
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include <new>
#include <stddef.h>

static thread_local char g_err[512] = "";

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

// The error path of the library's other sources: sets gam_last_error()'s
// message ("msg: detail") and returns code.
__attribute__((visibility("hidden"))) int gam_fail_(int code, const char *msg, const char *detail)
{
    return fail(code, detail && detail[0] ? "%s: %s" : "%s", msg, detail);
}

#define HIPCHK(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(GAM_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_));    \
    } while (0)

enum { KG_PREDICT = 0, KG_MOM2FACE, KG_EMF, KG_FLUX, KG_UPDATE, KG_BC, KG_BFLD, KG_SPARE };
static const char *kg_names[GAM_KG_COUNT] = {"predict", "mom2face", "emf", "flux",
                                             "update_dt", "bcs", "bflux2fld_ghosts", "unused"};

struct gam_ctx {
    gam_dims dims;
    gam_model model;
    MParm mp;
    Geo g;
    hipStream_t stream;
    double *st[2];        // State / oState buffers (S_NPL planes each)
    int cur;              // which of st[] is State
    unsigned gen;         // bumped whenever recorded launch arguments change
    double *H;            // half state + primitive planes
    double *met;          // metric planes
    double *bg;           // background planes (+9 scratch)
    double *gflx, *mflx, *vf;
    double *stage;        // one-plane staging buffer for layout conversion
    double *part;         // reduction partials
    Clock *clk;
    double *wind;         // device wind series
    WindDev wd;
    int haveGrid, haveBG, haveWind;
    // profiling
    int prof;
    hipEvent_t ev[2 * GAM_KG_COUNT * 8];
    int nev;
    double kg_ms[GAM_KG_COUNT];
    int kg_n[GAM_KG_COUNT];
    std::vector<int> pend;   // group id of each pending event pair
    // launch plan and its tuning
    Plan plan;
    int tune_slot;           // slot whose launches get events, or -1
    int ntev;
    hipEvent_t tev[2 * GAM_TUNE_MAXREPS];
    // two steps captured as one graph, from State buffer 0 / 1 (null: not
    // captured yet or dropped because a launch argument changed)
    hipGraphExec_t gx[2];
    // Fork/join of the independent flux and EMF launches inside captured
    // step graphs: two more streams (default priority) and the events that
    // order them; fork = 1 while a capture may branch.
    int concurrent;          // branch the captured graphs (default 1)
    int fork;
    hipStream_t aux[2];
    hipEvent_t fev[6];
    // set by the adjoint's forward steps with the dt sensitivities through
    // CalcDT: record each step's CalcDT argmin into dtrec (DT_REC doubles)
    double *dtrec, *dtrec_part;
    // bumped by everything that may write the forward intermediates (H, vf,
    // gflx, mflx): steps, step and ensemble graph launches, the adjoint's
    // recomputes
    unsigned long itgen;
    // when set, the next step writes its new State there instead of over
    // oState's buffer, and that buffer becomes State (the adjoint's interval
    // recompute steps through its tape this way)
    double *snx;
};

static void graphs_drop(gam_ctx *c)
{
    c->gen++;   // ensembles recorded with the old arguments are stale
    if (!c->gx[0] && !c->gx[1]) return;
    hipStreamSynchronize(c->stream);   // no replay may be in flight
    for (int q = 0; q < 2; q++)
        if (c->gx[q]) {
            hipGraphExecDestroy(c->gx[q]);
            c->gx[q] = nullptr;
        }
}

// ------------------------------------------------------------ geometry
static Box mkbox(int i0, int i1, int j0, int j1, int k0, int k1)
{
    Box b = {i0, i1, j0, j1, k0, k1};
    return b;
}
static Box face_box(const Geo &g) { return mkbox(g.isg, g.ieg + 1, g.jsg, g.jeg + 1, g.ksg, g.keg + 1); }
static Box cell_box(const Geo &g) { return mkbox(g.isg, g.ieg, g.jsg, g.jeg, g.ksg, g.keg); }
static Box act_box(const Geo &g) { return mkbox(g.is, g.ie, g.js, g.je, g.ks, g.ke); }
static Box dt_box(const Geo &g) { return mkbox(g.is - 1, g.ie + 1, g.js - 1, g.je + 1, g.ks - 1, g.ke + 1); }
static void edge_boxes(const Geo &g, Box *e)
{   // active EMF edges (fields.F90:100-175)
    e[0] = mkbox(g.is, g.ie, g.js, g.je + 1, g.ks, g.ke + 1);
    e[1] = mkbox(g.is, g.ie + 1, g.js, g.je, g.ks, g.ke + 1);
    e[2] = mkbox(g.is, g.ie + 1, g.js, g.je + 1, g.ks, g.ke);
}

static double *S_cur(gam_ctx *c) { return c->st[c->cur]; }
static double *S_old(gam_ctx *c) { return c->st[1 - c->cur]; }

// ------------------------------------------------------------ profiling
static int prof_flush(gam_ctx *c);
static const int PROF_CAP = (int)(sizeof(((gam_ctx *)0)->ev) / sizeof(hipEvent_t));
static void prof_begin(gam_ctx *c, int grp)
{
    if (!c->prof) return;
    if (c->nev + 2 > PROF_CAP) prof_flush(c);
    hipEventRecord(c->ev[c->nev], c->stream);
    c->pend.push_back(grp);
}
static void prof_end(gam_ctx *c, int grp)
{
    (void)grp;
    if (!c->prof) return;
    hipEventRecord(c->ev[c->nev + 1], c->stream);
    c->nev += 2;
}
static int prof_flush(gam_ctx *c)
{
    if (!c->prof || c->nev == 0) return GAM_OK;
    HIPCHK(hipEventSynchronize(c->ev[c->nev - 1]));
    for (int q = 0; q < c->nev / 2; q++) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev[2 * q], c->ev[2 * q + 1]));
        c->kg_ms[c->pend[q]] += ms;
        c->kg_n[c->pend[q]] += 1;
    }
    c->nev = 0;
    c->pend.clear();
    return GAM_OK;
}

// ------------------------------------------------------------ create
extern "C" const char *gam_last_error(void) { return g_err; }

extern "C" int gam_create(gam_ctx **out, const gam_dims *d, const gam_model *m)
{
    if (!out || !d || !m) return fail(GAM_E_ARG, "gam_create: NULL argument");
    *out = nullptr;
    if (d->nG != 4) return fail(GAM_E_ARG, "gam_create: nG must be 4 (got %d)", d->nG);
    if (d->Nip < 16) return fail(GAM_E_ARG, "Grid I-size smaller than VECLEN! (Nip=%d < 16)", d->Nip);
    if (d->Nip % 8 != 0) return fail(GAM_E_ARG, "Grid divisibility error, Grid I-size not divisible by 8 (Nip=%d)", d->Nip);
    if (d->Njp < 8 || d->Nkp < 8) return fail(GAM_E_ARG, "gam_create: Njp and Nkp must be >= 8 (got %d, %d)", d->Njp, d->Nkp);
    for (int s = 0; s < 6; s++) {
        int k = m->bc[s];
        if (k < GAM_BC_PERIODIC || k > GAM_BC_WIND) return fail(GAM_E_ARG, "gam_create: bad BC kind %d on side %d", k, s);
        if (k == GAM_BC_SPHIN && s != 0) return fail(GAM_E_ARG, "gam_create: sphere-in BC only on the inner I face");
        if ((k == GAM_BC_SPHOUT || k == GAM_BC_WIND) && s != 1) return fail(GAM_E_ARG, "gam_create: sphere-out/wind BC only on the outer I face");
    }
    if (m->rmeth != GAM_RECON_8CENT && m->rmeth != GAM_RECON_7UP) return fail(GAM_E_ARG, "Uknown sim/rmeth (%d)", m->rmeth);
    if (!(m->gamma > 1.0)) return fail(GAM_E_ARG, "gam_create: gamma must exceed 1");

    gam_ctx *c = (gam_ctx *)calloc(1, sizeof(gam_ctx));
    if (!c) return fail(GAM_E_NOMEM, "gam_create: host allocation failed");
    new (&c->pend) std::vector<int>();
    c->dims = *d;
    c->model = *m;
    MParm &p = c->mp;
    p.gamma = m->gamma; p.CFL = m->CFL; p.Vd0 = m->Vd0; p.Ca = m->Ca; p.cHogH = m->cHogH;
    p.cHogM = m->cHogM; p.pdmb = m->pdmb; p.dFloor = m->dFloor; p.pFloor = m->pFloor;
    p.tFin = m->tFin; p.dt0 = m->dt0; p.limDT0 = m->limDT0;
    p.igm1 = 1.0 / (m->gamma - 1.0);
    p.iCa2 = 1.0 / (m->Ca * m->Ca);
    p.doMHD = m->doMHD; p.doBoris = m->doBoris; p.doBG = m->doBackground; p.doLFM = m->doLFMLim;
    p.rmeth = m->rmeth;

    Geo &g = c->g;
    g.Nip = d->Nip; g.Njp = d->Njp; g.Nkp = d->Nkp;
    g.Ni = d->Nip + 8; g.Nj = d->Njp + 8; g.Nk = d->Nkp + 8;
    g.is = 1; g.ie = d->Nip; g.js = 1; g.je = d->Njp; g.ks = 1; g.ke = d->Nkp;
    g.isg = -3; g.ieg = g.ie + 4; g.jsg = -3; g.jeg = g.je + 4; g.ksg = -3; g.keg = g.ke + 4;
    g.P = ((g.Ni + 1 + 4) + 7) / 8 * 8;
    g.PP = (long)g.P * (g.Nj + 1);
    // i=isg lands at offset 4 of its row, so i=is (=isg+4) is 64-byte aligned
    g.c0 = 4 - g.isg - (long)g.P * g.jsg - g.PP * g.ksg;
    long n3 = g.PP * (g.Nk + 1);
    g.N3 = (n3 + 63) / 64 * 64;
    const size_t pl = (size_t)g.N3 * sizeof(double);

    int rc = GAM_OK;
    hipError_t e = hipSuccess;
#define ALLOC(ptr, nplanes)                                                   \
    if (e == hipSuccess) {                                                    \
        e = hipMalloc((void **)&(ptr), (size_t)(nplanes) * pl);              \
        if (e == hipSuccess) e = hipMemset((ptr), 0, (size_t)(nplanes) * pl); \
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int q = 0; e == hipSuccess && q < 2; q++) e = hipStreamCreateWithFlags(&c->aux[q], hipStreamNonBlocking);
    for (int q = 0; e == hipSuccess && q < 6; q++) e = hipEventCreateWithFlags(&c->fev[q], hipEventDisableTiming);
    {
        const char *cv = getenv("GAM_CONCURRENT");
        c->concurrent = (cv && cv[0] == '0') ? 0 : 1;
    }
    ALLOC(c->st[0], S_NPL);
    ALLOC(c->st[1], S_NPL);
    ALLOC(c->H, H_NPL);
    ALLOC(c->met, M_NPL);
    ALLOC(c->bg, B_NPL + 9);
    ALLOC(c->gflx, 15);
    ALLOC(c->mflx, 9);
    ALLOC(c->vf, 9);
    ALLOC(c->stage, 1);
    ALLOC(c->part, 1);
#undef ALLOC
    if (e == hipSuccess) e = hipMalloc((void **)&c->clk, sizeof(Clock));
    if (e == hipSuccess) {
        Clock ck;
        memset(&ck, 0, sizeof(ck));
        double h = GAM_HUGE;
        memcpy(&ck.dtBits, &h, sizeof(double));
        e = hipMemcpy(c->clk, &ck, sizeof(Clock), hipMemcpyHostToDevice);
    }
    for (int q = 0; e == hipSuccess && q < (int)(sizeof(c->ev) / sizeof(c->ev[0])); q++)
        e = hipEventCreate(&c->ev[q]);
    if (e != hipSuccess) {
        rc = fail(GAM_E_HIP, "gam_create: %s", hipGetErrorString(e));
        gam_destroy(c);
        return rc;
    }
    for (int q = 0; e == hipSuccess && q < 2 * GAM_TUNE_MAXREPS; q++) e = hipEventCreate(&c->tev[q]);
    if (e != hipSuccess) {
        rc = fail(GAM_E_HIP, "gam_create: %s", hipGetErrorString(e));
        gam_destroy(c);
        return rc;
    }
    c->plan = default_plan();
    c->tune_slot = -1;
    c->cur = 0;
    *out = c;
    return GAM_OK;
}

extern "C" int gam_destroy(gam_ctx *c)
{
    if (!c) return GAM_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    graphs_drop(c);
    double *bufs[] = {c->st[0], c->st[1], c->H, c->met, c->bg, c->gflx, c->mflx, c->vf, c->stage, c->part, c->wind};
    for (double *b : bufs) if (b) hipFree(b);
    if (c->clk) hipFree(c->clk);
    for (int q = 0; q < (int)(sizeof(c->ev) / sizeof(c->ev[0])); q++) if (c->ev[q]) hipEventDestroy(c->ev[q]);
    for (int q = 0; q < 2 * GAM_TUNE_MAXREPS; q++) if (c->tev[q]) hipEventDestroy(c->tev[q]);
    for (int q = 0; q < 2; q++) if (c->aux[q]) hipStreamDestroy(c->aux[q]);
    for (int q = 0; q < 6; q++) if (c->fev[q]) hipEventDestroy(c->fev[q]);
    if (c->stream) hipStreamDestroy(c->stream);
    c->pend.~vector<int>();
    free(c);
    return GAM_OK;
}

// ------------------------------------------------------------ transfers
// Upload ncomp consecutive Fortran planes of (ni,nj,nk) into device planes.
// The Fortran arrays may sit in host memory or in device memory of the
// context's GPU (hipMemcpyDefault lets the runtime pick the direction), so
// a caller holding its state on the GPU (torch tensors) moves nothing over
// PCIe.
static int upload(gam_ctx *c, const double *src, int face, int ncomp, double *dst)
{
    const Geo &g = c->g;
    const int ni = g.Ni + face, nj = g.Nj + face, nk = g.Nk + face;
    const size_t n = (size_t)ni * nj * nk;
    for (int q = 0; q < ncomp; q++) {
        HIPCHK(hipMemcpyAsync(c->stage, src + q * n, n * sizeof(double), hipMemcpyDefault, c->stream));
        launch_repack(c->stream, 1, g, ni, nj, nk, c->stage, dst + (size_t)q * g.N3);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return GAM_OK;
}

static int download(gam_ctx *c, const double *src, int face, int ncomp, double *dst)
{
    const Geo &g = c->g;
    const int ni = g.Ni + face, nj = g.Nj + face, nk = g.Nk + face;
    const size_t n = (size_t)ni * nj * nk;
    for (int q = 0; q < ncomp; q++) {
        launch_repack(c->stream, 0, g, ni, nj, nk, src + (size_t)q * g.N3, c->stage);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(dst + q * n, c->stage, n * sizeof(double), hipMemcpyDefault, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return GAM_OK;
}

extern "C" int gam_set_grid(gam_ctx *c, const gam_grid_arrays *a)
{
    if (!c || !a) return fail(GAM_E_ARG, "gam_set_grid: NULL argument");
    graphs_drop(c);   // launch arguments change
    if (!a->xyzcc || !a->xfc || !a->di || !a->dj || !a->dk || !a->volume || !a->Tf || !a->Te ||
        !a->Teb || !a->face || !a->edge)
        return fail(GAM_E_ARG, "gam_set_grid: every metric array is required");
    const long N3 = c->g.N3;
    int rc;
    if ((rc = upload(c, a->xyzcc, 0, 3, c->met + M_XYZCC * N3))) return rc;
    if ((rc = upload(c, a->xfc, 1, 9, c->met + M_XFC * N3))) return rc;
    if ((rc = upload(c, a->di, 0, 1, c->met + M_DI * N3))) return rc;
    if ((rc = upload(c, a->dj, 0, 1, c->met + M_DJ * N3))) return rc;
    if ((rc = upload(c, a->dk, 0, 1, c->met + M_DK * N3))) return rc;
    if ((rc = upload(c, a->volume, 0, 1, c->met + M_VOL * N3))) return rc;
    if ((rc = upload(c, a->Tf, 1, 27, c->met + M_TF * N3))) return rc;
    if ((rc = upload(c, a->Te, 1, 27, c->met + M_TE * N3))) return rc;
    if ((rc = upload(c, a->Teb, 1, 12, c->met + M_TEB * N3))) return rc;
    if ((rc = upload(c, a->face, 1, 3, c->met + M_FACE * N3))) return rc;
    if ((rc = upload(c, a->edge, 1, 3, c->met + M_EDGE * N3))) return rc;
    c->haveGrid = 1;
    return GAM_OK;
}

extern "C" int gam_grid_from_corners(gam_ctx *c, const double *xyz)
{
    if (!c || !xyz) return fail(GAM_E_ARG, "gam_grid_from_corners: NULL argument");
    graphs_drop(c);   // launch arguments change
    int rc = upload(c, xyz, 1, 3, c->met + (long)M_XYZ * c->g.N3);
    if (rc) return rc;
    launch_grid_from_corners(c->stream, c->g, c->met);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->haveGrid = 1;
    return GAM_OK;
}

extern "C" int gam_get_grid(gam_ctx *c, const gam_grid_arrays *a)
{
    if (!c || !a) return fail(GAM_E_ARG, "gam_get_grid: NULL argument");
    const long N3 = c->g.N3;
    int rc = GAM_OK;
    if (a->xyzcc && (rc = download(c, c->met + M_XYZCC * N3, 0, 3, a->xyzcc))) return rc;
    if (a->xfc && (rc = download(c, c->met + M_XFC * N3, 1, 9, a->xfc))) return rc;
    if (a->di && (rc = download(c, c->met + M_DI * N3, 0, 1, a->di))) return rc;
    if (a->dj && (rc = download(c, c->met + M_DJ * N3, 0, 1, a->dj))) return rc;
    if (a->dk && (rc = download(c, c->met + M_DK * N3, 0, 1, a->dk))) return rc;
    if (a->volume && (rc = download(c, c->met + M_VOL * N3, 0, 1, a->volume))) return rc;
    if (a->Tf && (rc = download(c, c->met + M_TF * N3, 1, 27, a->Tf))) return rc;
    if (a->Te && (rc = download(c, c->met + M_TE * N3, 1, 27, a->Te))) return rc;
    if (a->Teb && (rc = download(c, c->met + M_TEB * N3, 1, 12, a->Teb))) return rc;
    if (a->face && (rc = download(c, c->met + M_FACE * N3, 1, 3, a->face))) return rc;
    if (a->edge && (rc = download(c, c->met + M_EDGE * N3, 1, 3, a->edge))) return rc;
    return GAM_OK;
}

extern "C" int gam_set_background(gam_ctx *c, const gam_background_arrays *b)
{
    if (!c || !b) return fail(GAM_E_ARG, "gam_set_background: NULL argument");
    graphs_drop(c);   // launch arguments change
    if (!b->fcB0 || !b->edgB0 || !b->B0 || !b->bFlux0 || !b->dpB0)
        return fail(GAM_E_ARG, "gam_set_background: every background array is required");
    const long N3 = c->g.N3;
    int rc;
    if ((rc = upload(c, b->fcB0, 0, 9, c->bg + B_FCB0 * N3))) return rc;
    if ((rc = upload(c, b->edgB0, 0, 6, c->bg + B_EDGB0 * N3))) return rc;
    if ((rc = upload(c, b->B0, 0, 3, c->bg + B_B0 * N3))) return rc;
    if ((rc = upload(c, b->bFlux0, 0, 3, c->bg + B_BFLUX0 * N3))) return rc;
    if ((rc = upload(c, b->dpB0, 0, 3, c->bg + B_DPB0 * N3))) return rc;
    c->haveBG = 1;
    return GAM_OK;
}

extern "C" int gam_add_dipole_b0(gam_ctx *c, double M0)
{
    if (!c) return fail(GAM_E_ARG, "gam_add_dipole_b0: NULL context");
    if (!c->haveGrid) return fail(GAM_E_STATE, "gam_add_dipole_b0: grid metrics must be set first");
    graphs_drop(c);   // launch arguments change
    launch_dipole_b0(c->stream, c->g, M0, c->met, c->bg, c->bg + (long)B_NPL * c->g.N3);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->haveBG = 1;
    return GAM_OK;
}

extern "C" int gam_get_background(gam_ctx *c, const gam_background_arrays *b)
{
    if (!c || !b) return fail(GAM_E_ARG, "gam_get_background: NULL argument");
    const long N3 = c->g.N3;
    int rc = GAM_OK;
    if (b->fcB0 && (rc = download(c, c->bg + B_FCB0 * N3, 0, 9, b->fcB0))) return rc;
    if (b->edgB0 && (rc = download(c, c->bg + B_EDGB0 * N3, 0, 6, b->edgB0))) return rc;
    if (b->B0 && (rc = download(c, c->bg + B_B0 * N3, 0, 3, b->B0))) return rc;
    if (b->bFlux0 && (rc = download(c, c->bg + B_BFLUX0 * N3, 0, 3, b->bFlux0))) return rc;
    if (b->dpB0 && (rc = download(c, c->bg + B_DPB0 * N3, 0, 3, b->dpB0))) return rc;
    return GAM_OK;
}

extern "C" int gam_set_wind(gam_ctx *c, int n, const double *tW, const double *Q, const double *B,
                            double Bx0, double ByC, double BzC, const double *xyzW)
{
    if (!c || !tW || !Q || !B) return fail(GAM_E_ARG, "gam_set_wind: NULL argument");
    graphs_drop(c);   // launch arguments change
    if (n < 2) return fail(GAM_E_ARG, "gam_set_wind: need at least 2 samples (got %d)", n);
    // tMin/tMax/dtW as readWind (wind.F90:685-699)
    double tMin = tW[0], tMax = tW[0], dtMin = 1e300, dtMax = -1e300;
    for (int q = 0; q < n; q++) { tMin = fmin(tMin, tW[q]); tMax = fmax(tMax, tW[q]); }
    for (int q = 1; q < n; q++) { dtMin = fmin(dtMin, tW[q] - tW[q - 1]); dtMax = fmax(dtMax, tW[q] - tW[q - 1]); }
    if (!(dtMin > 0.0)) return fail(GAM_E_ARG, "gam_set_wind: times must increase");
    if (fabs(dtMin - dtMax) > 1.0e-6 * sqrt(dtMin * dtMin + dtMax * dtMax))
        return fail(GAM_E_ARG, "gam_set_wind: wind time series must be uniformly spaced");
    HIPCHK(hipStreamSynchronize(c->stream));   // no pending step may still read the old series
    // a series of the same length reuses the device buffer, so launches
    // recorded earlier (ensembles) never hold a freed address
    if (c->wind && c->wd.n != n) { hipFree(c->wind); c->wind = nullptr; }
    if (!c->wind) HIPCHK(hipMalloc((void **)&c->wind, sizeof(double) * (size_t)n * 9));
    std::vector<double> h((size_t)n * 9);
    memcpy(h.data(), tW, sizeof(double) * n);
    memcpy(h.data() + n, Q, sizeof(double) * n * 5);
    memcpy(h.data() + 6 * n, B, sizeof(double) * n * 3);
    HIPCHK(hipMemcpy(c->wind, h.data(), sizeof(double) * n * 9, hipMemcpyHostToDevice));
    WindDev &w = c->wd;
    w.tW = c->wind; w.Q = c->wind + n; w.B = c->wind + 6 * n; w.n = n;
    w.tMin = tMin; w.tMax = tMax; w.dtW = dtMin;
    w.Bx0 = Bx0; w.ByC = ByC; w.BzC = BzC;
    w.xW = xyzW ? xyzW[0] : 0.0; w.yW = xyzW ? xyzW[1] : 0.0; w.zW = xyzW ? xyzW[2] : 0.0;
    c->haveWind = 1;
    return GAM_OK;
}

static double *state_ptr(gam_ctx *c, int which)
{
    if (which == GAM_STATE) return S_cur(c);
    if (which == GAM_OSTATE) return S_old(c);
    return nullptr;
}

extern "C" int gam_put_state(gam_ctx *c, int which, const double *Gas, const double *magFlux,
                             const double *Bxyz, const double *Efld, double time)
{
    if (!c) return fail(GAM_E_ARG, "gam_put_state: NULL context");
    double *S = state_ptr(c, which);
    if (!S) return fail(GAM_E_ARG, "gam_put_state: which must be GAM_STATE or GAM_OSTATE");
    const long N3 = c->g.N3;
    int rc;
    if (Gas && (rc = upload(c, Gas, 0, 5, S + S_GAS * N3))) return rc;
    if (magFlux && (rc = upload(c, magFlux, 1, 3, S + S_MF * N3))) return rc;
    if (Bxyz && (rc = upload(c, Bxyz, 0, 3, S + S_B * N3))) return rc;
    if (Efld) {
        if ((rc = upload(c, Efld, 0, 3, S + S_E * N3))) return rc;
        Box eb[3];
        edge_boxes(c->g, eb);
        launch_mask_e(c->stream, c->g, cell_box(c->g), eb, S + S_E * N3);
        HIPCHK(hipGetLastError());
    }
    // State%time / oState%time
    const size_t off = (which == GAM_STATE) ? offsetof(Clock, st) : offsetof(Clock, ost);
    HIPCHK(hipMemcpyAsync((char *)c->clk + off, &time, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return GAM_OK;
}

extern "C" int gam_get_state(gam_ctx *c, int which, double *Gas, double *magFlux, double *Bxyz,
                             double *Efld, double *time)
{
    if (!c) return fail(GAM_E_ARG, "gam_get_state: NULL context");
    const long N3 = c->g.N3;
    double *S = state_ptr(c, which);
    int rc;
    if (which == GAM_HFSTATE) {
        if (Gas && (rc = download(c, c->H + H_GAS * N3, 0, 5, Gas))) return rc;
        if (magFlux && (rc = download(c, c->H + H_MF * N3, 1, 3, magFlux))) return rc;
        if (Bxyz && (rc = download(c, c->H + H_B * N3, 0, 3, Bxyz))) return rc;
        if (Efld) memset(Efld, 0, sizeof(double) * (size_t)c->g.Ni * c->g.Nj * c->g.Nk * 3);
        if (time) *time = 0.0;
        return GAM_OK;
    }
    if (!S) return fail(GAM_E_ARG, "gam_get_state: bad selector %d", which);
    if (Gas && (rc = download(c, S + S_GAS * N3, 0, 5, Gas))) return rc;
    if (magFlux && (rc = download(c, S + S_MF * N3, 1, 3, magFlux))) return rc;
    if (Bxyz && (rc = download(c, S + S_B * N3, 0, 3, Bxyz))) return rc;
    if (Efld && (rc = download(c, S + S_E * N3, 0, 3, Efld))) return rc;
    if (time) {
        Clock ck;
        HIPCHK(hipMemcpy(&ck, c->clk, sizeof(Clock), hipMemcpyDeviceToHost));
        *time = (which == GAM_STATE) ? ck.st : ck.ost;
    }
    return GAM_OK;
}

extern "C" int gam_get_solver(gam_ctx *c, int which, double *out)
{
    if (!c || !out) return fail(GAM_E_ARG, "gam_get_solver: NULL argument");
    if (which == 0) return download(c, c->gflx, 0, 15, out);
    if (which == 1) return download(c, c->mflx, 0, 9, out);
    return fail(GAM_E_ARG, "gam_get_solver: which must be 0 (gFlx) or 1 (mFlx)");
}

extern "C" int gam_set_time(gam_ctx *c, double t, double dt, int ts)
{
    if (!c) return fail(GAM_E_ARG, "gam_set_time: NULL context");
    HIPCHK(hipStreamSynchronize(c->stream));
    Clock ck;
    HIPCHK(hipMemcpy(&ck, c->clk, sizeof(Clock), hipMemcpyDeviceToHost));
    ck.t = t; ck.dt = dt; ck.dtA = dt; ck.ts = ts;
    HIPCHK(hipMemcpy(c->clk, &ck, sizeof(Clock), hipMemcpyHostToDevice));
    return GAM_OK;
}

extern "C" int gam_get_dims(gam_ctx *c, gam_dims *d)
{
    if (!c || !d) return fail(GAM_E_ARG, "gam_get_dims: NULL argument");
    *d = c->dims;
    return GAM_OK;
}

extern "C" int gam_set_dt0(gam_ctx *c, double dt0)
{
    if (!c) return fail(GAM_E_ARG, "gam_set_dt0: NULL context");
    if (!(dt0 >= 0.0)) return fail(GAM_E_ARG, "gam_set_dt0: dt0 must be >= 0");
    HIPCHK(hipStreamSynchronize(c->stream));
    c->mp.dt0 = dt0;
    c->model.dt0 = dt0;
    return GAM_OK;
}

extern "C" int gam_get_dt0(gam_ctx *c, double *dt0)
{
    if (!c || !dt0) return fail(GAM_E_ARG, "gam_get_dt0: NULL argument");
    *dt0 = c->mp.dt0;
    return GAM_OK;
}

extern "C" int gam_get_time(gam_ctx *c, double *t, double *dt, int *ts)
{
    if (!c) return fail(GAM_E_ARG, "gam_get_time: NULL context");
    Clock ck;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&ck, c->clk, sizeof(Clock), hipMemcpyDeviceToHost));
    if (t) *t = ck.t;
    if (dt) *dt = ck.dt;
    if (ts) *ts = ck.ts;
    return GAM_OK;
}

// ------------------------------------------------------------ step pieces
static int check_ready(gam_ctx *c, const char *who)
{
    if (!c) return fail(GAM_E_ARG, "%s: NULL context", who);
    if (!c->haveGrid) return fail(GAM_E_STATE, "%s: grid metrics not set", who);
    if (c->mp.doBG && !c->haveBG) return fail(GAM_E_STATE, "%s: doBackground but no background field", who);
    for (int s = 0; s < 6; s++)
        if (c->model.bc[s] == GAM_BC_WIND && !c->haveWind) return fail(GAM_E_STATE, "%s: wind BC without a wind series", who);
    return GAM_OK;
}

static Box empty_box() { return mkbox(1, 0, 1, 0, 1, 0); }

// bFlux2Fld over all cells, or over the ghost cells only when the active
// cells' faces are known to be unchanged since their B was computed.
static void enq_bflux2fld(gam_ctx *c, double *S, bool ghosts_only = false)
{
    const long N3 = c->g.N3;
    if (ghosts_only)
        launch_bflux2fld_ghosts(c->stream, c->g, S + S_MF * N3, S + S_B * N3, c->met);
    else
        launch_bflux2fld(c->stream, c->g, cell_box(c->g), empty_box(), S + S_MF * N3, S + S_B * N3, c->met);
}

// ghostB: whether BCs also set ghost-cell B (EnforceBCs on its own does;
// inside stepGamera the bFlux2Fld that follows rewrites every ghost cell).
// clk: the clock whose time the inflow BC reads (the context's by default).
static void enq_bcs(gam_ctx *c, double *S, bool ghostB = true, const Clock *clk = nullptr)
{
    for (int ax = 0; ax < 3; ax++) {
        const int lo = 2 * ax, hi = lo + 1;
        if (c->model.bc[lo] == GAM_BC_PERIODIC && c->model.bc[hi] == GAM_BC_PERIODIC) {
            launch_bc_periodic_pair(c->stream, ax, c->g, S);
            continue;
        }
        for (int side = lo; side <= hi; side++)
            launch_bc(c->stream, side, c->model.bc[side], c->g, c->mp, clk ? clk : c->clk, &c->wd, S, c->met,
                      ghostB);
    }
}

// While a slot is being tuned, its launches are bracketed by events.
static void tune_mark(gam_ctx *c, int slot)
{
    if (c->tune_slot != slot || c->ntev >= 2 * GAM_TUNE_MAXREPS) return;
    hipEventRecord(c->tev[c->ntev++], c->stream);
}

// AdvanceMHD on the device (mhdgroup.F90:16-76); swaps State/oState.
static void enq_advance(gam_ctx *c, int finmode)
{
    c->itgen++;
    const Geo &g = c->g;
    const long N3 = g.N3;
    hipStream_t s = c->stream;
    double *S = S_cur(c), *O = S_old(c);
    double *Sn = c->snx ? c->snx : O;   // becomes State

    prof_begin(c, KG_PREDICT);
    tune_mark(c, TS_PREDICT);
    launch_predict(s, c->plan.v[TS_PREDICT], g, c->mp, face_box(g), c->clk, S, O, c->met, c->H);
    tune_mark(c, TS_PREDICT);
    prof_end(c, KG_PREDICT);

    // interface fluxes (CalcStress), with Mom2Face of the same faces; the
    // EMFs do not depend on the fluxes, so they run after them.  Inside a
    // captured graph (fork) the j and k fluxes and the EMFs that take their
    // face velocities run on two more streams: the graph then has three
    // branches between the predictor and the update, which the device may
    // overlap.  The launches and their arguments are the same either way.
    const bool fork = c->fork && c->mp.doMHD;
    hipStream_t fs[3] = {s, s, s};   // stream of the flux of direction d and of its EMF
    if (fork) {
        order_record(s, c->fev[0], 0);
        for (int q = 0; q < 2; q++) {
            order_wait(c->aux[q], c->fev[0], 0);
            fs[1 + q] = c->aux[q];
        }
    }
    for (int d = 0; d < 3; d++) {
        Box fb = act_box(g);
        if (d == 0) fb.i1 += 1; else if (d == 1) fb.j1 += 1; else fb.k1 += 1;
        prof_begin(c, KG_FLUX);
        tune_mark(c, TS_FLUX_I + d);
        launch_flux(fs[d], d, c->plan.v[TS_FLUX_I + d], g, c->mp, fb, c->H, S, c->met, c->bg, c->gflx, c->mflx,
                    c->vf + (long)3 * d * N3);
        tune_mark(c, TS_FLUX_I + d);
        prof_end(c, KG_FLUX);
    }
    Box eb[3];
    edge_boxes(g, eb);
    if (c->mp.doMHD) {
        // Mom2Face on the rest of the ranges the EMF stencils read
        // (fields.F90:119-175)
        prof_begin(c, KG_MOM2FACE);
        launch_vface_rim(s, g, c->H, c->met, c->vf);
        prof_end(c, KG_MOM2FACE);
        if (fork) {
            order_record(s, c->fev[1], 1);
            for (int q = 0; q < 2; q++) order_wait(c->aux[q], c->fev[1], 1);
        }
        for (int eD = 0; eD < 3; eD++) {
            // edge direction eD takes the face velocities along dT1
            const int dT1 = (eD == 0) ? 1 : (eD == 1) ? 2 : 0;
            prof_begin(c, KG_EMF);
            tune_mark(c, TS_EMF_I + eD);
            launch_emf(fs[dT1], eD, c->plan.v[TS_EMF_I + eD], g, c->mp, eb[eD], c->clk, c->H, S,
                       c->vf + (long)3 * dT1 * N3, c->met, c->bg, Sn + S_E * N3);
            tune_mark(c, TS_EMF_I + eD);
            prof_end(c, KG_EMF);
        }
        if (fork)
            for (int q = 0; q < 2; q++) {
                order_record(c->aux[q], c->fev[2 + q], 2 + q);
                order_wait(s, c->fev[2 + q], 2 + q);
            }
    }

    prof_begin(c, KG_UPDATE);
    Box ctb = mkbox(g.is, g.ie + 1, g.js, g.je + 1, g.ks, g.ke + 1);
    tune_mark(c, TS_UPDATE);
    launch_update(s, c->plan.v[TS_UPDATE], g, c->mp, face_box(g), ctb, eb, act_box(g), dt_box(g), c->clk,
                  S, Sn, c->gflx, c->mflx, c->met, c->bg);
    tune_mark(c, TS_UPDATE);
    prof_end(c, KG_UPDATE);

    launch_finalize(s, c->mp, c->clk, finmode);
    if (c->snx) {
        c->st[1 - c->cur] = c->snx;
        c->snx = nullptr;
    }
    c->cur = 1 - c->cur;
}

// The forward intermediates of one step as enq_advance computes them
// (predictor, interface fluxes with their face velocities, the Mom2Face
// rim), without the rest of the step.
static void enq_intermediates(gam_ctx *c)
{
    c->itgen++;
    const Geo &g = c->g;
    const long N3 = g.N3;
    hipStream_t s = c->stream;
    double *S = S_cur(c), *O = S_old(c);
    launch_predict(s, c->plan.v[TS_PREDICT], g, c->mp, face_box(g), c->clk, S, O, c->met, c->H);
    for (int d = 0; d < 3; d++) {
        Box fb = act_box(g);
        if (d == 0) fb.i1 += 1; else if (d == 1) fb.j1 += 1; else fb.k1 += 1;
        launch_flux(s, d, c->plan.v[TS_FLUX_I + d], g, c->mp, fb, c->H, S, c->met, c->bg, c->gflx, c->mflx,
                    c->vf + (long)3 * d * N3);
    }
    if (c->mp.doMHD) launch_vface_rim(s, g, c->H, c->met, c->vf);
}

static void enq_step(gam_ctx *c)
{
    enq_advance(c, FIN_ADVANCE | FIN_MODEL | FIN_CALCDT | FIN_SETDT);
    double *S = S_cur(c);
    if (c->dtrec)
        launch_dt_record(c->stream, c->g, c->mp, dt_box(c->g), S, c->met, c->bg, c->clk, c->dtrec_part, c->dtrec);
    prof_begin(c, KG_BC);
    enq_bcs(c, S, false);
    prof_end(c, KG_BC);
    prof_begin(c, KG_BFLD);
    enq_bflux2fld(c, S, true);
    prof_end(c, KG_BFLD);
}

// n steps: pairs as one launch of a graph of two steps (captured on first
// use for the current State buffer; a step is only kernels, so the graph
// replays exactly what enq_step launches), the odd one plainly.  While
// profiling or tuning every step is enqueued plainly (those add events).
static int enq_steps(gam_ctx *c, int nsteps)
{
    int n = 0;
    const bool graph = !c->prof && c->tune_slot < 0;
    for (; graph && n + 2 <= nsteps; n += 2) {
        const int p = c->cur;
        if (!c->gx[p]) {
            hipGraph_t gr = nullptr;
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            c->fork = c->concurrent;
            enq_step(c);
            enq_step(c);
            c->fork = 0;
            hipError_t e = hipStreamEndCapture(c->stream, &gr);
            c->cur = p;
            if (e == hipSuccess) e = hipGraphInstantiate(&c->gx[p], gr, nullptr, nullptr, 0);
            if (gr) hipGraphDestroy(gr);
            if (e != hipSuccess) {
                c->gx[p] = nullptr;
                return fail(GAM_E_HIP, "step graph capture: %s", hipGetErrorString(e));
            }
        }
        HIPCHK(hipGraphLaunch(c->gx[p], c->stream));   // two steps: the parity comes back to p
        c->itgen++;
    }
    for (; n < nsteps; n++) enq_step(c);
    return GAM_OK;
}

static int read_clock(gam_ctx *c, Clock *ck)
{
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(ck, c->clk, sizeof(Clock), hipMemcpyDeviceToHost));
    return GAM_OK;
}

static int status_code(int st)
{
    if (st >= 2) return fail(GAM_E_DT_DISASTER, "<Timestep too small or no timestep below HUGE>");
    if (st == 1) return fail(GAM_E_DT_BAD, "<Drop in timestep by 10x or below limDT0>");
    return GAM_OK;
}

extern "C" int gam_init_run(gam_ctx *c, double *dt)
{
    int rc = check_ready(c, "gam_init_run");
    if (rc) return rc;
    double *S = S_cur(c), *O = S_old(c);
    const size_t bytes = (size_t)S_NPL * c->g.N3 * sizeof(double);
    if (c->mp.doMHD) enq_bflux2fld(c, S);
    enq_bcs(c, S);
    HIPCHK(hipMemcpyAsync(O, S, bytes, hipMemcpyDeviceToDevice, c->stream));
    launch_calc_dt(c->stream, c->g, c->mp, dt_box(c->g), c->clk, S, c->met, c->bg);
    // oState = State, then oState%time = State%time - dt (init.F90:228-244)
    Clock ck;
    launch_finalize(c->stream, c->mp, c->clk, FIN_CALCDT | FIN_SETDT | FIN_INIT);
    HIPCHK(hipGetLastError());
    if ((rc = read_clock(c, &ck))) return rc;
    if (dt) *dt = ck.dt;
    return status_code(ck.lastStatus);
}

extern "C" int gam_advance_mhd(gam_ctx *c, double dt)
{
    int rc = check_ready(c, "gam_advance_mhd");
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync((char *)c->clk + offsetof(Clock, dtA), &dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    enq_advance(c, FIN_ADVANCE);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return prof_flush(c);
}

extern "C" int gam_calc_dt(gam_ctx *c, double *dt)
{
    int rc = check_ready(c, "gam_calc_dt");
    if (rc) return rc;
    launch_calc_dt(c->stream, c->g, c->mp, dt_box(c->g), c->clk, S_cur(c), c->met, c->bg);
    launch_finalize(c->stream, c->mp, c->clk, FIN_CALCDT);
    HIPCHK(hipGetLastError());
    Clock ck;
    if ((rc = read_clock(c, &ck))) return rc;
    if (dt) *dt = ck.dtRet;
    return status_code(ck.lastStatus);
}

extern "C" int gam_enforce_bcs(gam_ctx *c)
{
    int rc = check_ready(c, "gam_enforce_bcs");
    if (rc) return rc;
    enq_bcs(c, S_cur(c));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return GAM_OK;
}

extern "C" int gam_bflux2fld(gam_ctx *c, int which)
{
    int rc = check_ready(c, "gam_bflux2fld");
    if (rc) return rc;
    double *S = state_ptr(c, which);
    if (!S) return fail(GAM_E_ARG, "gam_bflux2fld: which must be GAM_STATE or GAM_OSTATE");
    enq_bflux2fld(c, S);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return GAM_OK;
}

extern "C" int gam_step(gam_ctx *c, double *dt_next)
{
    int rc = check_ready(c, "gam_step");
    if (rc) return rc;
    enq_step(c);
    HIPCHK(hipGetLastError());
    if (!dt_next) return GAM_OK;
    Clock ck;
    if ((rc = read_clock(c, &ck))) return rc;
    if ((rc = prof_flush(c))) return rc;
    *dt_next = ck.dt;
    return status_code(ck.lastStatus);
}

extern "C" int gam_run(gam_ctx *c, int nsteps)
{
    int rc = check_ready(c, "gam_run");
    if (rc) return rc;
    if (nsteps < 0) return fail(GAM_E_ARG, "gam_run: nsteps < 0");
    if ((rc = enq_steps(c, nsteps))) return rc;
    HIPCHK(hipGetLastError());
    Clock ck;
    if ((rc = read_clock(c, &ck))) return rc;
    if ((rc = prof_flush(c))) return rc;
    return status_code(ck.status);
}

// gamAdvanceModel (gamtypessub.F90:74-90): one stepGamera, then more while
// Model%t < t0 + dt.  The loop condition needs Model%t, so every step ends
// with a 64-byte clock read; a CalcDT failure ends the loop with its code,
// where the reference prints and stops (step.F90:136-151).
extern "C" int gam_advance_model(gam_ctx *c, double dt, int *nsteps)
{
    int rc = check_ready(c, "gam_advance_model");
    if (rc) return rc;
    if (!(dt >= 0.0)) return fail(GAM_E_ARG, "gam_advance_model: dt must be >= 0");
    Clock ck;
    if ((rc = read_clock(c, &ck))) return rc;
    const double target = ck.t + dt;
    int n = 0;
    do {
        enq_step(c);
        HIPCHK(hipGetLastError());
        n++;
        if ((rc = read_clock(c, &ck))) break;
        if ((rc = status_code(ck.lastStatus))) break;
    } while (ck.t < target);
    if (nsteps) *nsteps = n;
    if (rc) return rc;
    return prof_flush(c);
}

extern "C" int gam_sync(gam_ctx *c)
{
    if (!c) return fail(GAM_E_ARG, "gam_sync: NULL context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return prof_flush(c);
}

extern "C" int gam_divb(gam_ctx *c, double *tot)
{
    int rc = check_ready(c, "gam_divb");
    if (rc) return rc;
    if (!tot) return fail(GAM_E_ARG, "gam_divb: NULL output");
    int np = 0;
    launch_divb(c->stream, c->g, act_box(c->g), S_cur(c) + S_MF * c->g.N3, c->bg, c->mp.doBG, c->part, &np);
    HIPCHK(hipGetLastError());
    std::vector<double> h(np);
    HIPCHK(hipMemcpyAsync(h.data(), c->part, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (int q = 0; q < np; q++) s += h[q];
    *tot = s;
    return GAM_OK;
}

extern "C" int gam_time_steps(gam_ctx *c, int nsteps, double *ms)
{
    int rc = check_ready(c, "gam_time_steps");
    if (rc) return rc;
    if (nsteps < 1 || !ms) return fail(GAM_E_ARG, "gam_time_steps: need nsteps >= 1 and an output");
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipEventRecord(a, c->stream));
    if ((rc = enq_steps(c, nsteps))) return rc;
    HIPCHK(hipEventRecord(b, c->stream));
    HIPCHK(hipEventSynchronize(b));
    float f = 0.0f;
    HIPCHK(hipEventElapsedTime(&f, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    *ms = f;
    Clock ck;
    if ((rc = read_clock(c, &ck))) return rc;
    if ((rc = prof_flush(c))) return rc;
    return status_code(ck.status);
}

extern "C" int gam_set_profiling(gam_ctx *c, int on)
{
    if (!c) return fail(GAM_E_ARG, "gam_set_profiling: NULL context");
    int rc = gam_sync(c);
    if (rc) return rc;
    c->prof = on ? 1 : 0;
    for (int q = 0; q < GAM_KG_COUNT; q++) { c->kg_ms[q] = 0.0; c->kg_n[q] = 0; }
    c->nev = 0;
    c->pend.clear();
    return GAM_OK;
}

extern "C" int gam_kernel_times(gam_ctx *c, double *ms, int *counts)
{
    if (!c) return fail(GAM_E_ARG, "gam_kernel_times: NULL context");
    int rc = gam_sync(c);
    if (rc) return rc;
    for (int q = 0; q < GAM_KG_COUNT; q++) {
        if (ms) ms[q] = c->kg_ms[q];
        if (counts) counts[q] = c->kg_n[q];
    }
    return GAM_OK;
}

extern "C" const char *gam_kernel_group_name(int grp)
{
    if (grp < 0 || grp >= GAM_KG_COUNT) return "";
    return kg_names[grp];
}

// ------------------------------------------------------------ ensembles
// Every member's step is recorded (not issued) as its list of step-kernel
// launches, twice: once from each of its State/oState roles.  Launch p of
// all members becomes one batched launch (k_ens) over the members'
// concatenated grids with a device table of their argument records; the
// batched launches of one step are captured as a hipGraph per role, with
// the members' fork/join of the flux and EMF branches (the same streams
// and events at the same places, so the graph has the same branches).
struct EnsPhase {
    void (*launch)(hipStream_t, const void *, const void *, unsigned, dim3);
    dim3 block;
    unsigned nwg;
    size_t hdr_off, rec_off;   // byte offsets into gam_ens::dev
    int lane;                  // 0 the main stream, 1 and 2 the branch streams
    int mark, ev;              // EnsLaunch::mark and ::ev; mark 0 is a launch
};

struct gam_ens {
    int n;
    gam_ctx **m;
    int *cur0;                       // each member's State buffer at recording
    unsigned *gen0;                  // each member's launch-argument generation at recording
    int parity;                      // steps taken mod 2 since recording
    hipStream_t stream;
    hipStream_t aux[2];              // branch streams of the capture
    hipEvent_t ev[4];                // fork/join events of the capture
    unsigned char *dev;              // headers and argument records
    std::vector<EnsPhase> ph[2];
    hipGraph_t graph[2];
    hipGraphExec_t exec[2];
};

extern "C" int gam_ens_destroy(gam_ens *e)
{
    if (!e) return GAM_OK;
    if (e->stream) hipStreamSynchronize(e->stream);
    for (int q = 0; q < 2; q++) {
        if (e->exec[q]) hipGraphExecDestroy(e->exec[q]);
        if (e->graph[q]) hipGraphDestroy(e->graph[q]);
    }
    if (e->dev) hipFree(e->dev);
    if (e->stream) hipStreamDestroy(e->stream);
    for (int q = 0; q < 2; q++) if (e->aux[q]) hipStreamDestroy(e->aux[q]);
    for (int q = 0; q < 4; q++) if (e->ev[q]) hipEventDestroy(e->ev[q]);
    free(e->m);
    free(e->cur0);
    free(e->gen0);
    e->ph[0].~vector<EnsPhase>();
    e->ph[1].~vector<EnsPhase>();
    free(e);
    return GAM_OK;
}

static bool same_variant(const gam_ctx *a, const gam_ctx *b)
{
    const gam_model &x = a->model, &y = b->model;
    if (x.rmeth != y.rmeth || x.doMHD != y.doMHD || x.doBoris != y.doBoris ||
        x.doBackground != y.doBackground || x.doLFMLim != y.doLFMLim)
        return false;
    for (int s = 0; s < 6; s++) if (x.bc[s] != y.bc[s]) return false;
    return true;
}

extern "C" int gam_ens_create(gam_ens **out, int n, gam_ctx *const *members)
{
    if (!out || !members) return fail(GAM_E_ARG, "gam_ens_create: NULL argument");
    *out = nullptr;
    if (n < 1 || n > GAM_ENS_MAX) return fail(GAM_E_ARG, "gam_ens_create: n = %d outside 1..%d", n, GAM_ENS_MAX);
    int rc;
    for (int q = 0; q < n; q++) {
        if ((rc = check_ready(members[q], "gam_ens_create"))) return rc;
        if (!same_variant(members[0], members[q]))
            return fail(GAM_E_ARG, "gam_ens_create: member %d differs from member 0 in rmeth, doMHD, doBoris, "
                                   "doBackground, doLFMLim or BC kinds", q);
        for (int r = 0; r < q; r++)
            if (members[r] == members[q]) return fail(GAM_E_ARG, "gam_ens_create: member %d repeats member %d", q, r);
    }
    gam_ens *e = (gam_ens *)calloc(1, sizeof(gam_ens));
    if (!e) return fail(GAM_E_NOMEM, "gam_ens_create: host allocation failed");
    new (&e->ph[0]) std::vector<EnsPhase>();
    new (&e->ph[1]) std::vector<EnsPhase>();
    e->n = n;
    e->m = (gam_ctx **)calloc(n, sizeof(gam_ctx *));
    e->cur0 = (int *)calloc(n, sizeof(int));
    e->gen0 = (unsigned *)calloc(n, sizeof(unsigned));
    if (!e->m || !e->cur0 || !e->gen0) {
        gam_ens_destroy(e);
        return fail(GAM_E_NOMEM, "gam_ens_create: host allocation failed");
    }

    // record two steps of every member (enq_step swaps its buffers twice)
    std::vector<std::vector<EnsLaunch>> rec[2];
    rec[0].resize(n);
    rec[1].resize(n);
    for (int q = 0; q < n; q++) {
        gam_ctx *c = members[q];
        e->m[q] = c;
        e->cur0[q] = c->cur;
        e->gen0[q] = c->gen;
        const int prof = c->prof;
        c->prof = 0;
        c->fork = c->concurrent;
        for (int r = 0; r < 2; r++) {
            set_launch_recorder(&rec[r][q]);
            enq_step(c);
        }
        set_launch_recorder(nullptr);
        c->fork = 0;
        c->prof = prof;
    }
    // every member issues the same number of launches per step; at each
    // position the members are grouped by kernel (their launch plans may
    // pick different shapes), one batched launch per group
    auto lane_of = [&](int q, hipStream_t st) {
        return st == members[q]->aux[0] ? 1 : st == members[q]->aux[1] ? 2 : 0;
    };
    for (int r = 0; r < 2; r++)
        for (int q = 1; q < n; q++) {
            bool same = rec[r][q].size() == rec[r][0].size();
            for (size_t p = 0; same && p < rec[r][0].size(); p++) {
                const EnsLaunch &A = rec[r][0][p], &B = rec[r][q][p];
                same = A.mark == B.mark && A.ev == B.ev && lane_of(0, A.stream) == lane_of(q, B.stream);
            }
            if (!same) {
                gam_ens_destroy(e);
                return fail(GAM_E_ARG, "gam_ens_create: member %d issues a different kernel sequence", q);
            }
        }
    // device layout: per batched launch an EnsHdr, its owner table and the
    // group's argument records, each 256-aligned
    std::vector<unsigned char> host;
    auto put = [&](const void *p, size_t bytes) -> size_t {
        const size_t off = (host.size() + 255) / 256 * 256;
        host.resize(off + bytes);
        memcpy(host.data() + off, p, bytes);
        return off;
    };
    for (int r = 0; r < 2; r++) {
        for (size_t p = 0; p < rec[r][0].size(); p++) {
            const int lane = lane_of(0, rec[r][0][p].stream);
            if (rec[r][0][p].mark) {
                EnsPhase ph = {};
                ph.lane = lane;
                ph.mark = rec[r][0][p].mark;
                ph.ev = rec[r][0][p].ev;
                e->ph[r].push_back(ph);
                continue;
            }
            std::vector<std::vector<int>> groups;
            for (int q = 0; q < n; q++) {
                const EnsLaunch &L = rec[r][q][p];
                size_t gi = 0;
                for (; gi < groups.size(); gi++) {
                    const EnsLaunch &F = rec[r][groups[gi][0]][p];
                    if (F.launch == L.launch && F.block.x == L.block.x && F.block.y == L.block.y &&
                        F.block.z == L.block.z && F.rec.size() == L.rec.size())
                        break;
                }
                if (gi == groups.size()) groups.emplace_back();
                groups[gi].push_back(q);
            }
            for (const std::vector<int> &grp : groups) {
                const int gn = (int)grp.size();
                EnsHdr h;
                memset(&h, 0, sizeof(h));
                h.n = (unsigned)gn;
                unsigned long tot = 0;
                for (int q = 0; q < gn; q++) {
                    const EnsLaunch &L = rec[r][grp[q]][p];
                    h.wg0[q] = (unsigned)tot;
                    h.gx[q] = L.grid.x;
                    h.gy[q] = L.grid.y;
                    h.gz[q] = L.grid.z;
                    tot += ((unsigned long)L.grid.x * L.grid.y * L.grid.z + 7) / 8 * 8;
                }
                if (tot > 0x7fffffffUL) {
                    gam_ens_destroy(e);
                    return fail(GAM_E_ARG, "gam_ens_create: batched grid too large");
                }
                h.wg0[gn] = (unsigned)tot;
                std::vector<unsigned short> own(tot / 8);
                for (int q = 0; q < gn; q++)
                    for (unsigned w = h.wg0[q]; w < h.wg0[q + 1]; w += 8) own[w / 8] = (unsigned short)q;
                const EnsLaunch &F = rec[r][grp[0]][p];
                EnsPhase ph = {};
                ph.lane = lane;
                ph.launch = F.launch;
                ph.block = F.block;
                ph.nwg = (unsigned)tot;
                ph.hdr_off = put(&h, sizeof(h));
                host.resize(ph.hdr_off + sizeof(h) + own.size() * sizeof(unsigned short));
                memcpy(host.data() + ph.hdr_off + sizeof(h), own.data(), own.size() * sizeof(unsigned short));
                const size_t rs = F.rec.size();
                ph.rec_off = put(F.rec.data(), rs);
                for (int q = 1; q < gn; q++) {
                    const size_t at = host.size();
                    host.resize(at + rs);
                    memcpy(host.data() + at, rec[r][grp[q]][p].rec.data(), rs);
                }
                e->ph[r].push_back(ph);
            }
        }
    }
    hipError_t err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    for (int q = 0; err == hipSuccess && q < 2; q++) err = hipStreamCreateWithFlags(&e->aux[q], hipStreamNonBlocking);
    for (int q = 0; err == hipSuccess && q < 4; q++) err = hipEventCreateWithFlags(&e->ev[q], hipEventDisableTiming);
    if (err == hipSuccess) err = hipMalloc((void **)&e->dev, host.size());
    if (err == hipSuccess) err = hipMemcpy(e->dev, host.data(), host.size(), hipMemcpyHostToDevice);
    for (int r = 0; err == hipSuccess && r < 2; r++) {
        err = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
        if (err != hipSuccess) break;
        const hipStream_t ls[3] = {e->stream, e->aux[0], e->aux[1]};
        for (const EnsPhase &ph : e->ph[r]) {
            if (ph.mark == 1) err = hipEventRecord(e->ev[ph.ev], ls[ph.lane]);
            else if (ph.mark == 2) err = hipStreamWaitEvent(ls[ph.lane], e->ev[ph.ev], 0);
            else ph.launch(ls[ph.lane], e->dev + ph.hdr_off, e->dev + ph.rec_off, ph.nwg, ph.block);
            if (err != hipSuccess) break;
        }
        const hipError_t ec = hipStreamEndCapture(e->stream, &e->graph[r]);
        if (err == hipSuccess) err = ec;
        if (err == hipSuccess) err = hipGraphInstantiate(&e->exec[r], e->graph[r], nullptr, nullptr, 0);
    }
    if (err != hipSuccess) {
        rc = fail(GAM_E_HIP, "gam_ens_create: %s", hipGetErrorString(err));
        gam_ens_destroy(e);
        return rc;
    }
    *out = e;
    return GAM_OK;
}

// Issue nsteps graph launches; members' buffers and parity follow.
static int ens_enqueue(gam_ens *e, int nsteps)
{
    int q0 = -1;
    for (int q = 0; q < e->n; q++) {
        gam_ctx *c = e->m[q];
        if (c->gen != e->gen0[q])
            return fail(GAM_E_STATE, "gam_ens_run: member %d's grid, background, inflow series or launch plan "
                        "changed since gam_ens_create (create the ensemble again)", q);
        const int rel = c->cur ^ e->cur0[q];
        if (q0 < 0) q0 = rel;
        if (rel != q0)
            return fail(GAM_E_STATE, "gam_ens_run: member %d was stepped an odd number of times on its own", q);
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    e->parity = q0;
    for (int n = 0; n < nsteps; n++) {
        HIPCHK(hipGraphLaunch(e->exec[e->parity], e->stream));
        e->parity ^= 1;
    }
    for (int q = 0; q < e->n; q++) e->m[q]->itgen++;
    for (int q = 0; q < e->n; q++) e->m[q]->cur = e->cur0[q] ^ e->parity;
    return GAM_OK;
}

static int ens_status(gam_ens *e)
{
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int q = 0; q < e->n; q++) {
        Clock ck;
        HIPCHK(hipMemcpy(&ck, e->m[q]->clk, sizeof(Clock), hipMemcpyDeviceToHost));
        if (ck.status) {
            status_code(ck.status);
            char msg[512];
            snprintf(msg, sizeof(msg), "%s", g_err);
            return fail(ck.status >= 2 ? GAM_E_DT_DISASTER : GAM_E_DT_BAD, "ensemble member %d: %s", q, msg);
        }
    }
    return GAM_OK;
}

extern "C" int gam_ens_run(gam_ens *e, int nsteps)
{
    if (!e) return fail(GAM_E_ARG, "gam_ens_run: NULL ensemble");
    if (nsteps < 0) return fail(GAM_E_ARG, "gam_ens_run: nsteps < 0");
    int rc = ens_enqueue(e, nsteps);
    if (rc) return rc;
    return ens_status(e);
}

extern "C" int gam_ens_time_steps(gam_ens *e, int nsteps, double *ms)
{
    if (!e) return fail(GAM_E_ARG, "gam_ens_time_steps: NULL ensemble");
    if (nsteps < 1 || !ms) return fail(GAM_E_ARG, "gam_ens_time_steps: need nsteps >= 1 and an output");
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    int rc = GAM_OK;
    for (int q = 0; q < e->n; q++) HIPCHK(hipStreamSynchronize(e->m[q]->stream));
    HIPCHK(hipEventRecord(a, e->stream));
    rc = ens_enqueue(e, nsteps);
    HIPCHK(hipEventRecord(b, e->stream));
    HIPCHK(hipEventSynchronize(b));
    float f = 0.0f;
    HIPCHK(hipEventElapsedTime(&f, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    *ms = f;
    if (rc) return rc;
    return ens_status(e);
}

extern "C" int gam_ens_launches_per_step(gam_ens *e, int *n)
{
    if (!e || !n) return fail(GAM_E_ARG, "gam_ens_launches_per_step: NULL argument");
    int k = 0;
    for (const EnsPhase &ph : e->ph[0]) k += ph.mark == 0;
    *n = k;
    return GAM_OK;
}

// ------------------------------------------------------------ launch plans
static const char *slot_names[GAM_TUNE_SLOTS] = {"predict", "emf_i", "emf_j", "emf_k",
                                                 "flux_i", "flux_j", "flux_k", "update"};
static_assert(GAM_TUNE_SLOTS == TS_COUNT, "slot count");

extern "C" const char *gam_tune_slot_name(int slot)
{
    if (slot < 0 || slot >= GAM_TUNE_SLOTS) return "";
    return slot_names[slot];
}

extern "C" int gam_tune_candidates(int slot, int *values)
{
    int v[GAM_TUNE_MAXCAND];
    const int n = plan_candidates(slot, v);
    if (values) for (int q = 0; q < n; q++) values[q] = v[q];
    return n;
}

extern "C" int gam_get_plan(gam_ctx *c, int *plan)
{
    if (!c || !plan) return fail(GAM_E_ARG, "gam_get_plan: NULL argument");
    for (int q = 0; q < TS_COUNT; q++) plan[q] = c->plan.v[q];
    return GAM_OK;
}

extern "C" int gam_set_plan(gam_ctx *c, const int *plan)
{
    if (!c || !plan) return fail(GAM_E_ARG, "gam_set_plan: NULL argument");
    graphs_drop(c);   // launch arguments change
    for (int q = 0; q < TS_COUNT; q++) {
        int v[GAM_TUNE_MAXCAND];
        const int n = plan_candidates(q, v);
        bool ok = false;
        for (int r = 0; r < n; r++) ok = ok || v[r] == plan[q];
        if (!ok) return fail(GAM_E_ARG, "gam_set_plan: %d is not a shape of slot %s", plan[q], slot_names[q]);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int q = 0; q < TS_COUNT; q++) c->plan.v[q] = plan[q];
    return GAM_OK;
}

// Time one slot's launches over reps steps (after one untimed step).
static int tune_time(gam_ctx *c, int slot, int reps, double *ms)
{
    c->tune_slot = -1;
    enq_step(c);
    c->tune_slot = slot;
    c->ntev = 0;
    for (int r = 0; r < reps; r++) enq_step(c);
    c->tune_slot = -1;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (int q = 0; q + 1 < c->ntev; q += 2) {
        float f = 0.0f;
        HIPCHK(hipEventElapsedTime(&f, c->tev[q], c->tev[q + 1]));
        tot += f;
    }
    *ms = tot;
    return GAM_OK;
}

// Time n whole steps as runs take them (step graphs, with their branches),
// after two untimed steps that capture the graph.
static int tune_whole(gam_ctx *c, int n, double *ms)
{
    graphs_drop(c);
    int rc = enq_steps(c, 2);
    if (rc) return rc;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, c->stream));
    rc = enq_steps(c, n);
    HIPCHK(hipEventRecord(b, c->stream));
    HIPCHK(hipEventSynchronize(b));
    float f = 0.0f;
    HIPCHK(hipEventElapsedTime(&f, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    *ms = f;
    return rc;
}

extern "C" int gam_autotune(gam_ctx *c, int reps)
{
    int rc = check_ready(c, "gam_autotune");
    if (rc) return rc;
    graphs_drop(c);   // launch arguments change
    if (reps < 1 || reps > GAM_TUNE_MAXREPS) return fail(GAM_E_ARG, "gam_autotune: reps must be in 1..%d", GAM_TUNE_MAXREPS);
    // snapshot State, oState and the clock; tuning takes real steps
    const size_t bytes = (size_t)S_NPL * c->g.N3 * sizeof(double);
    double *save = nullptr;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMalloc((void **)&save, 2 * bytes));
    Clock ck;
    const int cur0 = c->cur;
    const int prof = c->prof;
    hipError_t e = hipMemcpyAsync(save, c->st[0], bytes, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(save + bytes / sizeof(double), c->st[1], bytes, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(&ck, c->clk, sizeof(Clock), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        hipFree(save);
        return fail(GAM_E_HIP, "gam_autotune: %s", hipGetErrorString(e));
    }
    c->prof = 0;
    const Plan start = c->plan;
    // the current shape is timed first and another replaces it only when
    // more than 1 % faster, so timing noise does not flip near-equal shapes
    for (int slot = 0; slot < TS_COUNT && rc == GAM_OK; slot++) {
        int v[GAM_TUNE_MAXCAND];
        const int n = plan_candidates(slot, v);
        int bestv = c->plan.v[slot];
        double best = 0.0;
        rc = tune_time(c, slot, reps, &best);
        for (int q = 0; q < n && rc == GAM_OK; q++) {
            if (v[q] == bestv) continue;
            c->plan.v[slot] = v[q];
            double ms = 0.0;
            rc = tune_time(c, slot, reps, &ms);
            if (rc == GAM_OK && ms < 0.99 * best) { best = ms; bestv = v[q]; }
        }
        c->plan.v[slot] = bestv;
    }
    // Each slot was timed alone in plain steps; runs replay step graphs
    // whose flux and EMF branches overlap, so a changed plan replaces the
    // starting one only if whole graph steps are more than 1 % faster too.
    bool changed = false;
    for (int q = 0; q < TS_COUNT; q++) changed = changed || c->plan.v[q] != start.v[q];
    if (rc == GAM_OK && changed) {
        const Plan tuned = c->plan;
        double tn = 0.0, to = 0.0;
        rc = tune_whole(c, 4 * reps, &tn);
        c->plan = start;
        if (rc == GAM_OK) rc = tune_whole(c, 4 * reps, &to);
        graphs_drop(c);
        c->plan = (rc == GAM_OK && tn < 0.99 * to) ? tuned : start;
    }
    c->prof = prof;
    c->tune_slot = -1;
    // put the run back where it was
    hipStreamSynchronize(c->stream);
    e = hipMemcpyAsync(c->st[0], save, bytes, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->st[1], save + bytes / sizeof(double), bytes, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->clk, &ck, sizeof(Clock), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    c->cur = cur0;
    hipFree(save);
    if (rc) return rc;
    if (e != hipSuccess) return fail(GAM_E_HIP, "gam_autotune: %s", hipGetErrorString(e));
    return GAM_OK;
}

// ------------------------------------------------------------ adjoint
// Discrete adjoint of the step (include/gamadj.h).  The forward run keeps
// State/oState checkpoints every K steps and the clock of every step; the
// reverse sweep recomputes each checkpoint interval forward (keeping its
// states on a tape) and then runs the adjoint kernels of its steps from
// last to first.
enum { ADJ_SET = 11 };   // planes of an adjoint state set: gas 5, face flux 3, cell B 3
// planes a checkpoint or tape state keeps: gas, face flux, cell B (a step
// computes its edge EMF before reading it, so a state's E is never input)
enum { CK_NPL = S_E };
// planes of one step's forward intermediates: half state (H_NPL), face
// velocities 9, gas fluxes 15, magnetic fluxes 9
enum { IT_VF = H_NPL, IT_GF = IT_VF + 9, IT_MF = IT_GF + 15, IT_NPL = IT_MF + 9 };

// One step's forward intermediates (the context's own buffers, or a slot
// of the interval tape).
struct StepIT {
    double *H, *vf, *gflx, *mflx;
};

struct gam_adj {
    gam_ctx *c;
    int nsteps, K, nck;
    int cur0;                // State buffer at the start of the run
    int haveW, recorded;
    int lastTaped;           // the tapes hold the last interval from gam_adj_forward
    int ownLast;             // and the context's own intermediates its last step,
    unsigned long ownGen;    // while the context's itgen is still ownGen
    double *ck;              // nck x (State, oState), CK_NPL planes each
    double *fin;             // State, oState at the end
    Clock *clkh;             // clock before step n, n = 0..nsteps
    Clock *oclk;             // clock at the start with t = oState%time
    double *tape;            // K + 2 State buffers of one interval
    double *itape;           // Kt steps' forward intermediates (IT_NPL planes each), or null
    int Kt;                  // steps of an interval whose intermediates are taped (<= K)
    int ostAtState;          // start oState's ghosts follow the BCs at State%time
    // a second pair of interval tapes and a stream of its own: the reverse
    // sweep recomputes the next interval into one pair while the adjoint
    // steps of the current interval read the other (null: no overlap)
    double *tape2, *itape2;
    hipStream_t rs;
    hipEvent_t evR[2], evA[2];   // recompute of a pair done / adjoint done reading it
    double *lam;             // three adjoint state sets (ADJ_SET planes each)
    double *ag, *amf;        // adjoints of gFlx (15) and mFlx (9)
    double *awh, *awn, *ahg; // adjoints of H_WH, H_WN, H_GAS (5 each)
    double *ahmf, *ahb;      // adjoints of H_MF, H_B (3 each)
    double *avf, *aur, *adm, *ae;   // Vf (9), Reynolds gas (5), dM (3), E (3)
    double *rv;              // Mom2Face normalisation per face direction (3)
    double *fsc;             // flux-transpose scratch, 48 planes per face direction (MHD)
    double *esc;             // EMF-transpose scratch, 16 planes per edge direction (MHD);
                             // the first planes of fsc: the EMF gather has read it before
                             // the flux transpose writes fsc
    double *wgt;             // objective weights (1 plane)
    double *part;            // objective partial sums (1 plane)
    int nw;                  // inflow series length (0: no inflow BC)
    double *gwq, *gwb;       // adjoints of the inflow series Q (nw,5), B (nw,3), WIND_COPIES copies each
    hipEvent_t ev[4];
    float fwd_ms, bwd_ms;
    // dt sensitivities (gam_adj_set_dt_mode): 0 off, 1 dJ/d(dt_n) with the
    // other steps' dt held, 2 also through CalcDT (each dt follows the state),
    // 3 the first dt too (CalcDT of the start State)
    int dtMode, dtDone;
    double *mu;              // dJ/d(dt_n), n = 0..nsteps
    double *pdt;             // the transposes' per-point shares: pdt, then pw (2 planes)
    double *dtpart;          // block partial sums (2 per block of the face box)
    double *dts;             // a replay's dt sequence (nsteps)
    double *dtrec;           // each step's CalcDT argmin record (DT_REC doubles; mode >= 2),
                             // then the start State's (mode 3)
    double *dtrpart;         // block partials of the forward steps' argmin scans
    int dtrecMode;           // dt mode of the recorded forward
};

// Step n of a forward run records its CalcDT argmin (dt mode 2).
static void dt_rec_use(gam_adj *a, int n)
{
    gam_ctx *c = a->c;
    c->dtrec = a->dtMode >= 2 ? a->dtrec + (size_t)n * DT_REC : nullptr;
    c->dtrec_part = a->dtrpart;
}

static StepIT it_slot(gam_adj *a, int q, int buf = 0)
{
    const long N3 = a->c->g.N3;
    double *b = (buf ? a->itape2 : a->itape) + (size_t)q * IT_NPL * N3;
    return StepIT{b, b + (size_t)IT_VF * N3, b + (size_t)IT_GF * N3, b + (size_t)IT_MF * N3};
}

static bool it_taped(const gam_adj *a, int q) { return a->itape && q < a->Kt; }

// Point the context's forward intermediates at step q's tape slot (pair
// buf) when that step is taped, else at its own buffers.
static void it_use(gam_adj *a, int q, int buf, const StepIT &own)
{
    gam_ctx *c = a->c;
    const StepIT t = it_taped(a, q) ? it_slot(a, q, buf) : own;
    c->H = t.H; c->vf = t.vf; c->gflx = t.gflx; c->mflx = t.mflx;
}

extern "C" int gam_adj_destroy(gam_adj *a)
{
    if (!a) return GAM_OK;
    if (a->c && a->c->stream) hipStreamSynchronize(a->c->stream);
    double *bufs[] = {a->ck, a->fin, a->tape, a->itape, a->tape2, a->itape2, a->lam, a->ag, a->amf, a->awh, a->awn, a->ahg, a->ahmf,
                      a->ahb, a->avf, a->aur, a->adm, a->ae, a->wgt, a->part, a->gwq, a->gwb, a->rv,
                      a->fsc, a->mu, a->pdt, a->dtpart, a->dts, a->dtrec, a->dtrpart};
    for (double *b : bufs) if (b) hipFree(b);
    if (a->clkh) hipFree(a->clkh);
    if (a->oclk) hipFree(a->oclk);
    for (int q = 0; q < 4; q++) if (a->ev[q]) hipEventDestroy(a->ev[q]);
    for (int q = 0; q < 2; q++) {
        if (a->evR[q]) hipEventDestroy(a->evR[q]);
        if (a->evA[q]) hipEventDestroy(a->evA[q]);
    }
    if (a->rs) hipStreamDestroy(a->rs);
    free(a);
    return GAM_OK;
}

extern "C" int gam_adj_create(gam_adj **out, gam_ctx *c, int nsteps, int K)
{
    return gam_adj_create_budget(out, c, nsteps, K, -1, 1);
}

extern "C" int gam_adj_create_budget(gam_adj **out, gam_ctx *c, int nsteps, int K, int max_taped,
                                     int overlap)
{
    if (!out) return fail(GAM_E_ARG, "gam_adj_create: NULL argument");
    *out = nullptr;
    int rc = check_ready(c, "gam_adj_create");
    if (rc) return rc;
    if (!c->mp.doMHD && (c->mp.doBoris || c->mp.doBG))
        return fail(GAM_E_ARG, "gam_adj_create: Boris correction and background field need MHD");
    if (nsteps < 1 || K < 1 || K > nsteps) return fail(GAM_E_ARG, "gam_adj_create: need 1 <= ckpt_every <= nsteps");
    gam_adj *a = (gam_adj *)calloc(1, sizeof(gam_adj));
    if (!a) return fail(GAM_E_NOMEM, "gam_adj_create: host allocation failed");
    a->c = c;
    a->nsteps = nsteps;
    a->K = K;
    a->nck = (nsteps + K - 1) / K;
    const size_t pl = (size_t)c->g.N3 * sizeof(double);
    hipError_t e = hipSuccess;
#define AALLOC(ptr, bytes) if (e == hipSuccess) e = hipMalloc((void **)&(ptr), (bytes))
    AALLOC(a->ck, (size_t)a->nck * 2 * CK_NPL * pl);
    AALLOC(a->fin, 2 * S_NPL * pl);
    AALLOC(a->clkh, (size_t)(nsteps + 1) * sizeof(Clock));
    AALLOC(a->oclk, sizeof(Clock));
    AALLOC(a->tape, (size_t)(K + 2) * CK_NPL * pl);
    AALLOC(a->lam, 3 * ADJ_SET * pl);
    AALLOC(a->ag, 15 * pl);
    AALLOC(a->amf, 9 * pl);
    AALLOC(a->awh, 5 * pl);
    AALLOC(a->awn, 5 * pl);
    AALLOC(a->ahg, 5 * pl);
    AALLOC(a->ahmf, 3 * pl);
    AALLOC(a->ahb, 3 * pl);
    AALLOC(a->avf, 9 * pl);
    AALLOC(a->aur, 5 * pl);
    AALLOC(a->adm, 3 * pl);
    AALLOC(a->ae, 3 * pl);
    AALLOC(a->wgt, pl);
    AALLOC(a->part, pl);
    AALLOC(a->rv, 3 * pl);
    if (c->mp.doMHD) AALLOC(a->fsc, 3 * 48 * pl);
    a->esc = a->fsc;
    for (int sd = 0; sd < 6; sd++)
        if (c->model.bc[sd] == GAM_BC_WIND) a->nw = c->wd.n;
    if (a->nw) {
        AALLOC(a->gwq, (size_t)WIND_COPIES * a->nw * 5 * sizeof(double));
        AALLOC(a->gwb, (size_t)WIND_COPIES * a->nw * 3 * sizeof(double));
    }
#undef AALLOC
    // the interval's forward intermediates, kept where HBM allows (otherwise
    // each adjoint step recomputes its predictor, face velocities and fluxes)
    size_t mfree = 0, mtot = 0;
    if (e == hipSuccess) e = hipMemGetInfo(&mfree, &mtot);
    // (as many steps of an interval as fit, 4 GiB kept free)
    const size_t slot = (size_t)IT_NPL * pl, spare = (size_t)4 << 30;
    a->Kt = (e == hipSuccess && mfree > spare) ? (int)std::min<size_t>((size_t)K, (mfree - spare) / slot) : 0;
    if (max_taped >= 0) a->Kt = std::min(a->Kt, max_taped);
    const size_t itb = (size_t)a->Kt * slot;
    if (a->Kt > 0) {
        e = hipMalloc((void **)&a->itape, itb);
        if (e == hipSuccess) e = hipMemset(a->itape, 0, itb);
        // and, where HBM allows, the second pair for the overlapped sweep
        const size_t t2 = (size_t)(K + 2) * CK_NPL * pl;
        if (e == hipSuccess && overlap && a->Kt == K && a->nck > 1 && itb + t2 + spare < mfree - itb) {
            e = hipMalloc((void **)&a->itape2, itb);
            if (e == hipSuccess) e = hipMemset(a->itape2, 0, itb);
            if (e == hipSuccess) e = hipMalloc((void **)&a->tape2, t2);
            if (e == hipSuccess) e = hipStreamCreateWithFlags(&a->rs, hipStreamNonBlocking);
            for (int q = 0; e == hipSuccess && q < 2; q++) {
                e = hipEventCreateWithFlags(&a->evR[q], hipEventDisableTiming);
                if (e == hipSuccess) e = hipEventCreateWithFlags(&a->evA[q], hipEventDisableTiming);
            }
        }
    }
    if (e == hipSuccess) e = hipMemset(a->lam, 0, 3 * ADJ_SET * pl);
    if (e == hipSuccess) e = hipMemset(a->ag, 0, 15 * pl);
    if (e == hipSuccess) e = hipMemset(a->amf, 0, 9 * pl);
    if (e == hipSuccess) e = hipMemset(a->aur, 0, 5 * pl);
    if (e == hipSuccess) e = hipMemset(a->adm, 0, 3 * pl);
    if (e == hipSuccess) {
        launch_vface_rv(c->stream, c->g, c->met, a->rv);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    for (int q = 0; e == hipSuccess && q < 4; q++) e = hipEventCreate(&a->ev[q]);
    if (e != hipSuccess) {
        rc = fail(GAM_E_HIP, "gam_adj_create: %s", hipGetErrorString(e));
        gam_adj_destroy(a);
        return rc;
    }
    *out = a;
    return GAM_OK;
}

extern "C" int gam_adj_set_weights(gam_adj *a, const double *w)
{
    if (!a || !w) return fail(GAM_E_ARG, "gam_adj_set_weights: NULL argument");
    int rc = upload(a->c, w, 0, 1, a->wgt);
    if (rc) return rc;
    a->haveW = 1;
    return GAM_OK;
}

static int adj_objective(gam_adj *a, const double *S, double *lamS, double *J)
{
    gam_ctx *c = a->c;
    int np = 0;
    launch_adj_objective(c->stream, c->g, c->mp, S, a->wgt, lamS, a->part, &np);
    HIPCHK(hipGetLastError());
    std::vector<double> h(np);
    HIPCHK(hipMemcpyAsync(h.data(), a->part, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (int q = 0; q < np; q++) tot += h[q];
    if (J) *J = tot;
    return GAM_OK;
}

// Put the context at step n's recorded buffers: State/oState from src
// (two consecutive blocks of npl planes), clock from the record.
static int adj_put(gam_adj *a, int n, const double *src, int npl)
{
    gam_ctx *c = a->c;
    const size_t bytes = (size_t)npl * c->g.N3 * sizeof(double);
    c->cur = a->cur0 ^ (n & 1);
    HIPCHK(hipMemcpyAsync(S_cur(c), src, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(S_old(c), src + (size_t)npl * c->g.N3, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->clk, a->clkh + n, sizeof(Clock), hipMemcpyDeviceToDevice, c->stream));
    return GAM_OK;
}

extern "C" int gam_adj_forward(gam_adj *a, double *J)
{
    if (!a) return fail(GAM_E_ARG, "gam_adj_forward: NULL argument");
    if (!a->haveW) return fail(GAM_E_STATE, "gam_adj_forward: objective weights not set");
    gam_ctx *c = a->c;
    const long N3 = c->g.N3;
    const size_t bytes = (size_t)S_NPL * N3 * sizeof(double);
    const size_t ckb = (size_t)CK_NPL * N3 * sizeof(double);
    HIPCHK(hipStreamSynchronize(c->stream));
    a->cur0 = c->cur;
    const int prof = c->prof;
    c->prof = 0;
    HIPCHK(hipEventRecord(a->ev[0], c->stream));
    HIPCHK(hipMemcpyAsync(a->clkh, c->clk, sizeof(Clock), hipMemcpyDeviceToDevice, c->stream));
    // the last interval's states (and intermediates) go straight onto the
    // tapes, so the reverse sweep starts without recomputing it
    const int sL = (a->nck - 1) * a->K;
    // mode 3: the first dt is CalcDT of the start State as it is now
    if (a->dtMode == 3)
        launch_dt_record(c->stream, c->g, c->mp, dt_box(c->g), S_cur(c), c->met, c->bg, c->clk, a->dtrpart,
                         a->dtrec + (size_t)a->nsteps * DT_REC);
    const StepIT own = {c->H, c->vf, c->gflx, c->mflx};
    for (int n = 0; n < a->nsteps; n++) {
        if (n % a->K == 0) {
            double *ck = a->ck + (size_t)(n / a->K) * 2 * CK_NPL * N3;
            HIPCHK(hipMemcpyAsync(ck, S_cur(c), ckb, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(ck + (size_t)CK_NPL * N3, S_old(c), ckb, hipMemcpyDeviceToDevice, c->stream));
        }
        if (n >= sL) {
            const int q = n - sL;
            if (q == 0) HIPCHK(hipMemcpyAsync(a->tape, S_old(c), ckb, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(a->tape + (size_t)(1 + q) * CK_NPL * N3, S_cur(c), ckb,
                                  hipMemcpyDeviceToDevice, c->stream));
            it_use(a, q, 0, own);
        }
        dt_rec_use(a, n);
        enq_step(c);
        HIPCHK(hipMemcpyAsync(a->clkh + n + 1, c->clk, sizeof(Clock), hipMemcpyDeviceToDevice, c->stream));
    }
    c->dtrec = nullptr;
    a->dtrecMode = a->dtMode;
    c->H = own.H; c->vf = own.vf; c->gflx = own.gflx; c->mflx = own.mflx;
    HIPCHK(hipMemcpyAsync(a->tape + (size_t)(1 + a->nsteps - sL) * CK_NPL * N3, S_cur(c), ckb,
                          hipMemcpyDeviceToDevice, c->stream));
    a->lastTaped = 1;
    a->ownLast = !it_taped(a, a->nsteps - 1 - sL);
    a->ownGen = c->itgen;
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(a->ev[1], c->stream));
    c->prof = prof;
    HIPCHK(hipMemcpyAsync(a->fin, S_cur(c), bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(a->fin + (size_t)S_NPL * N3, S_old(c), bytes, hipMemcpyDeviceToDevice, c->stream));
    int rc = adj_objective(a, S_cur(c), a->lam, J);
    if (rc) return rc;
    HIPCHK(hipEventElapsedTime(&a->fwd_ms, a->ev[0], a->ev[1]));
    // the start oState's boundary values are those of its own time, or of
    // State's when the caller's forward starts from gam_init_run
    Clock ck;
    HIPCHK(hipMemcpy(&ck, a->clkh, sizeof(Clock), hipMemcpyDeviceToHost));
    if (!a->ostAtState) ck.t = ck.ost;
    HIPCHK(hipMemcpy(a->oclk, &ck, sizeof(Clock), hipMemcpyHostToDevice));
    if ((rc = read_clock(c, &ck))) return rc;
    a->recorded = 1;
    return status_code(ck.status);
}

static int adj_restore_final(gam_adj *a)
{
    gam_ctx *c = a->c;
    int rc = adj_put(a, a->nsteps, a->fin, S_NPL);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return GAM_OK;
}

// Transposes of the boundary pass on one adjoint set, sides OUTK to INI;
// S = the state the pass ran on (its active cells and faces), bclk = the
// clock it ran at.
static void adj_bcs(gam_adj *a, double *l, const double *S, const Clock *bclk)
{
    gam_ctx *c = a->c;
    for (int sd = 5; sd >= 0; sd--) {
        if (c->model.bc[sd] == GAM_BC_WIND)
            launch_bc_wind_adj(c->stream, c->g, c->mp, bclk, c->wd, S, l, c->met, a->gwq, a->gwb);
        else if (c->mp.doMHD) launch_bc_mhd_adj(c->stream, sd, c->model.bc[sd], c->g, l, c->met);
        else launch_bc_adj(c->stream, sd, c->model.bc[sd] == GAM_BC_PERIODIC ? 0 : 1, c->g, l);
    }
}

// One step's adjoint: lA (wrt the step's output State) and lB (wrt its
// output oState = input State) in, lB (input State) and lC (input oState)
// out.  Sn, On = the step's input State and oState, Sn1 its output State;
// clk its clock, bclk the clock after it (its boundary pass's time).
// it: the step's taped forward intermediates, or null to recompute them.
// n: the step's index (its dt sensitivity goes to mu[n], the predictor's
// share through odt to mu[n - 1]; with dt mode 2, mu[n + 1] goes back
// through the CalcDT that ends step n).
static void adj_step(gam_adj *a, int n, const StepIT *it, const double *Sn, const double *On, const double *Sn1,
                     const Clock *clk, const Clock *bclk, double *lA, double *lB, double *lC)
{
    gam_ctx *c = a->c;
    const Geo &g = c->g;
    const long N3 = g.N3;
    hipStream_t s = c->stream;
    const size_t pl = (size_t)N3 * sizeof(double);
    double *pdt = a->dtMode ? a->pdt : nullptr, *pw = a->dtMode ? a->pdt + N3 : nullptr;
    if (pdt) hipMemsetAsync(pdt, 0, 2 * pl, s);
    // CalcDT at the end of step n (before its boundary pass) gave dt[n + 1]
    auto dt_inject = [&]() {
        if (a->dtMode >= 2 && n + 1 < a->nsteps)
            launch_dt_inject(s, g, c->mp, a->dtrec + (size_t)n * DT_REC, a->mu + n + 1, lA);
    };
    Box eb[3];
    edge_boxes(g, eb);
    const double *H = it ? it->H : c->H, *vf = it ? it->vf : c->vf;
    const double *gflx = it ? it->gflx : c->gflx, *mflx = it ? it->mflx : c->mflx;
    // the forward intermediates the transposes read
    if (!it) {
        c->itgen++;
        launch_predict(s, c->plan.v[TS_PREDICT], g, c->mp, face_box(g), clk, Sn, On, c->met, c->H);
        // the hydrodynamic transposes read no face fluxes, but the dt share
        // of the update does
        if (c->mp.doMHD || pdt) {
            for (int d = 0; d < 3; d++) {
                Box fb = act_box(g);
                if (d == 0) fb.i1 += 1; else if (d == 1) fb.j1 += 1; else fb.k1 += 1;
                launch_flux(s, d, c->plan.v[TS_FLUX_I + d], g, c->mp, fb, c->H, Sn, c->met, c->bg, c->gflx,
                            c->mflx, c->vf + (long)3 * d * N3);
            }
            if (c->mp.doMHD) launch_vface_rim(s, g, c->H, c->met, c->vf);
        }
    }
    if (c->mp.doMHD) {
        // ghost bFlux2Fld, boundary pass, update (with its cell B)
        launch_cellb_adj(s, g, 0, lA, c->met);
        adj_bcs(a, lA, Sn1, bclk);
        dt_inject();
        const Box ctb = mkbox(g.is, g.ie + 1, g.js, g.je + 1, g.ks, g.ke + 1);
        launch_update_adj_mhd(s, g, c->mp, act_box(g), ctb, eb, clk, Sn, Sn1, gflx, mflx, c->met, c->bg,
                              lA, lB, a->aur, a->adm, a->ag, a->amf, a->ae, pdt);
        // the EMF transpose first: its gather writes aVf, aHMF and aHG whole,
        // the flux transpose adds into aHMF and aHG, and its gather writes
        // aWH, aWN and aHB whole (no plane needs zeroing)
        for (int ed = 0; ed < 3; ed++)
            launch_emf_adj(s, ed, g, c->mp, eb[ed], clk, H, Sn, vf, c->met, c->bg, a->ae, a->esc, pdt);
        launch_emf_adj_gather(s, g, c->mp, eb, a->esc, c->met, a->avf, a->ahmf, a->ahg, lB + S_MF * N3);
        for (int d = 0; d < 3; d++) {
            Box fb = act_box(g);
            if (d == 0) fb.i1 += 1; else if (d == 1) fb.j1 += 1; else fb.k1 += 1;
            launch_flux_adj_mhd(s, d, g, c->mp, fb, H, Sn, c->met, c->bg, a->ag, a->amf,
                                a->fsc + (size_t)48 * d * N3, a->ahmf, a->ahg);
        }
        launch_recon_adj_gather(s, g, c->mp, a->fsc, a->rv, a->avf, c->met, a->awh, a->awn, a->ahb,
                                lB + S_B * N3);
        launch_predict_adj_mhd(s, g, c->mp, clk, Sn, On, c->met, a->ahg, a->awh, a->awn, a->ahb, a->ahmf, lB, lC,
                               pw);
    } else {
        adj_bcs(a, lA, Sn1, bclk);
        dt_inject();
        launch_update_adj(s, g, act_box(g), clk, lA, lB, a->ag, c->met, gflx, pdt);
        hipMemsetAsync(a->awh, 0, 5 * pl, s);
        hipMemsetAsync(a->awn, 0, 5 * pl, s);
        for (int d = 0; d < 3; d++) {
            Box fb = act_box(g);
            if (d == 0) fb.i1 += 1; else if (d == 1) fb.j1 += 1; else fb.k1 += 1;
            launch_flux_adj(s, d, g, c->mp, fb, H, c->met, a->ag, a->awh, a->awn);
        }
        launch_predict_adj(s, g, c->mp, clk, Sn, On, c->met, a->awh, a->awn, lB, lC, pw);
    }
    if (pdt) launch_dt_fold(s, g, pdt, pw, a->dtpart, clk, n, a->mu);
}

extern "C" int gam_adj_backward(gam_adj *a, double *gGas, double *gGasO, double *gMF, double *gMFO)
{
    if (!a || !gGas) return fail(GAM_E_ARG, "gam_adj_backward: NULL argument");
    if (!a->recorded) return fail(GAM_E_STATE, "gam_adj_backward: no recorded forward run");
    gam_ctx *c = a->c;
    const Geo &g = c->g;
    const long N3 = g.N3;
    const size_t bytes = (size_t)CK_NPL * N3 * sizeof(double);   // a tape state
    hipStream_t s = c->stream;
    double *lA = a->lam, *lB = a->lam + ADJ_SET * N3, *lC = a->lam + 2 * ADJ_SET * N3;
    if (a->nw && a->nw != c->wd.n)
        return fail(GAM_E_STATE, "gam_adj_backward: the inflow series changed length since gam_adj_create");
    if (a->dtMode >= 2 && a->dtrecMode != a->dtMode)
        return fail(GAM_E_STATE, "gam_adj_backward: dt mode %d needs a gam_adj_forward run in that mode", a->dtMode);
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipEventRecord(a->ev[2], s));
    HIPCHK(hipMemsetAsync(a->lam, 0, 3 * ADJ_SET * N3 * sizeof(double), s));
    if (a->dtMode) HIPCHK(hipMemsetAsync(a->mu, 0, (size_t)(a->nsteps + 1) * sizeof(double), s));
    a->dtDone = 0;
    if (a->nw) {
        HIPCHK(hipMemsetAsync(a->gwq, 0, (size_t)WIND_COPIES * a->nw * 5 * sizeof(double), s));
        HIPCHK(hipMemsetAsync(a->gwb, 0, (size_t)WIND_COPIES * a->nw * 3 * sizeof(double), s));
    }
    int rc = adj_objective(a, a->fin, lA, nullptr);
    if (rc) return rc;
    const int prof = c->prof;
    c->prof = 0;
    // Interval seg recomputed from its checkpoint onto tape pair buf: on the
    // second stream when the sweep overlaps, else on the context's stream.
    const bool ovl = a->tape2 != nullptr;
    // Step whose forward intermediates the context's own buffers hold: the
    // last step of an interval recomputed without overlap, when it is not
    // taped.  Its adjoint step comes right after that recompute and uses
    // them instead of recomputing its predictor, face velocities and fluxes.
    int ownStep = -1;
    auto recompute = [&](int seg, int buf) -> int {
        const int s0 = seg * a->K, s1 = (s0 + a->K < a->nsteps) ? s0 + a->K : a->nsteps;
        const double *ck = a->ck + (size_t)seg * 2 * CK_NPL * N3;
        double *tp = buf ? a->tape2 : a->tape;
        const StepIT own = {c->H, c->vf, c->gflx, c->mflx};
        c->stream = ovl ? a->rs : s;
        // tape[0] = oState at s0, tape[1 + q] = State at step s0 + q, q = 0..s1-s0.
        // The steps run on the tape itself: step q reads State from slot
        // 1 + q and oState from slot q and writes its new State into slot
        // 2 + q (its 3 EMF planes reach past the slot's 11 into the first
        // planes of slot 3 + q, which nothing reads before the next step
        // overwrites them), so no State is copied onto the tape.
        auto slot = [&](int k) { return tp + (size_t)k * CK_NPL * N3; };
        double *keep[2] = {c->st[0], c->st[1]};
        c->cur = a->cur0 ^ (s0 & 1);
        int r = GAM_OK;
        if (hipMemcpyAsync(slot(1), ck, bytes, hipMemcpyDeviceToDevice, c->stream) ||
            hipMemcpyAsync(slot(0), ck + (size_t)CK_NPL * N3, bytes, hipMemcpyDeviceToDevice, c->stream) ||
            hipMemcpyAsync(c->clk, a->clkh + s0, sizeof(Clock), hipMemcpyDeviceToDevice, c->stream))
            r = fail(GAM_E_HIP, "gam_adj_backward: tape copy failed");
        c->st[c->cur] = slot(1);
        c->st[1 - c->cur] = slot(0);
        for (int q = 0; !r && q < s1 - s0; q++) {
            it_use(a, q, buf, own);   // a taped step writes its intermediates into its slot
            if (q + 1 < s1 - s0) {
                c->snx = slot(2 + q);
                dt_rec_use(a, s0 + q);
                enq_step(c);
            } else {
                // the last step: its adjoint step reads only its intermediates
                // (its dt record is the forward run's), and the State it
                // would give is the next checkpoint's, or the final one
                enq_intermediates(c);
            }
        }
        c->snx = nullptr;
        c->st[0] = keep[0];
        c->st[1] = keep[1];
        const double *nxt = seg + 1 < a->nck ? a->ck + (size_t)(seg + 1) * 2 * CK_NPL * N3 : a->fin;
        if (!r && hipMemcpyAsync(slot(1 + s1 - s0), nxt, bytes, hipMemcpyDeviceToDevice, c->stream))
            r = fail(GAM_E_HIP, "gam_adj_backward: tape copy failed");
        c->dtrec = nullptr;
        if (!r && ovl && hipEventRecord(a->evR[buf], c->stream)) r = fail(GAM_E_HIP, "gam_adj_backward: event");
        c->H = own.H; c->vf = own.vf; c->gflx = own.gflx; c->mflx = own.mflx;
        c->stream = s;
        ownStep = (!ovl && !it_taped(a, s1 - s0 - 1)) ? s1 - 1 : -1;
        return r;
    };
    // the last interval: from the recorded forward run's tapes, or recomputed
    const bool reuse = a->lastTaped;
    if (reuse && !ovl && a->ownLast && c->itgen == a->ownGen) ownStep = a->nsteps - 1;
    if (!reuse && (rc = recompute(a->nck - 1, 0))) return rc;
    for (int seg = a->nck - 1; seg >= 0; seg--) {
        const int s0 = seg * a->K, s1 = (s0 + a->K < a->nsteps) ? s0 + a->K : a->nsteps;
        const int buf = ovl ? (a->nck - 1 - seg) & 1 : 0;
        if (ovl) {
            // the next interval goes onto the other pair once the adjoint
            // steps that last read it are done
            if (seg > 0) {
                if (seg + 1 < a->nck) HIPCHK(hipStreamWaitEvent(a->rs, a->evA[buf ^ 1], 0));
                if ((rc = recompute(seg - 1, buf ^ 1))) return rc;
            }
            if (!(seg == a->nck - 1 && reuse)) HIPCHK(hipStreamWaitEvent(s, a->evR[buf], 0));
        } else if (seg < a->nck - 1 && (rc = recompute(seg, 0))) {
            return rc;
        }
        const double *tp = buf ? a->tape2 : a->tape;
        for (int n = s1 - 1; n >= s0; n--) {
            const int q = n - s0;
            const bool tpd = it_taped(a, q), own = !tpd && n == ownStep;
            const StepIT its = tpd ? it_slot(a, q, buf) : StepIT{c->H, c->vf, c->gflx, c->mflx};
            adj_step(a, n, (tpd || own) ? &its : nullptr, tp + (size_t)(1 + q) * CK_NPL * N3,
                     tp + (size_t)q * CK_NPL * N3, tp + (size_t)(2 + q) * CK_NPL * N3, a->clkh + n, a->clkh + n + 1,
                     lA, lB, lC);
            ownStep = -1;
            double *t = lA;
            lA = lB;
            lB = lC;
            lC = t;
        }
        if (ovl) HIPCHK(hipEventRecord(a->evA[buf], s));
        HIPCHK(hipGetLastError());
    }
    // pair 0 still holds the last interval only if no other went onto it
    a->lastTaped = (a->nck == 1);
    c->prof = prof;
    // mode 3: the first dt was CalcDT of the start State (after its boundary
    // pass): its share goes to that state before its ghosts are folded
    if (a->dtMode == 3) launch_dt_inject(s, g, c->mp, a->dtrec + (size_t)a->nsteps * DT_REC, a->mu, lA);
    // the start states' ghost cells follow from their interiors through the
    // boundary conditions, and their cell B from the face fluxes (in the
    // order of gam_adj_replay: active cell B, the boundary pass, every
    // cell's B): fold those adjoints into the interior gas and face fluxes.
    // The boundary pass reads active cells' B (the inflow's fast-speed fix,
    // the sphere walls' copies), so its transposes leave adjoints on active
    // cell B, which a second cell-B transpose folds into the face fluxes.
    if (c->mp.doMHD) {
        launch_cellb_adj(s, g, 1, lA, c->met);
        launch_cellb_adj(s, g, 1, lB, c->met);
    }
    adj_bcs(a, lA, a->ck, a->clkh);
    adj_bcs(a, lB, a->ck + (size_t)CK_NPL * N3, a->oclk);
    if (c->mp.doMHD) {
        launch_cellb_adj(s, g, 1, lA, c->met);
        launch_cellb_adj(s, g, 1, lB, c->met);
    }
    HIPCHK(hipEventRecord(a->ev[3], s));
    HIPCHK(hipEventSynchronize(a->ev[3]));
    HIPCHK(hipEventElapsedTime(&a->bwd_ms, a->ev[2], a->ev[3]));
    a->dtDone = a->dtMode;
    if ((rc = download(c, lA, 0, 5, gGas))) return rc;
    if (gGasO && (rc = download(c, lB, 0, 5, gGasO))) return rc;
    if (gMF && (rc = download(c, lA + S_MF * N3, 1, 3, gMF))) return rc;
    if (gMFO && (rc = download(c, lB + S_MF * N3, 1, 3, gMFO))) return rc;
    return adj_restore_final(a);
}

// Replay from the recorded start with the given start arrays; dt of each
// step: mode 0 the recorded one, 1 dts[n] (host array, nsteps values; 0
// takes CalcDT's dt: of the state the step before left, or for the first
// step of the start after its boundary pass, as gam_init_run does), 2 the
// recorded one for the first step and CalcDT's for the others.
static int adj_replay(gam_adj *a, const double *Gas, const double *oGas, const double *MF, const double *oMF,
                      int mode, const double *dts, double *J)
{
    gam_ctx *c = a->c;
    const long N3 = c->g.N3;
    if (mode == 1) {
        if (!a->dts) HIPCHK(hipMalloc((void **)&a->dts, (size_t)a->nsteps * sizeof(double)));
        HIPCHK(hipMemcpy(a->dts, dts, (size_t)a->nsteps * sizeof(double), hipMemcpyHostToDevice));
    }
    int rc = adj_put(a, 0, a->ck, CK_NPL);
    if (rc) return rc;
    if ((rc = upload(c, Gas, 0, 5, S_cur(c) + S_GAS * N3))) return rc;
    if (oGas && (rc = upload(c, oGas, 0, 5, S_old(c) + S_GAS * N3))) return rc;
    if (MF && (rc = upload(c, MF, 1, 3, S_cur(c) + S_MF * N3))) return rc;
    if (oMF && (rc = upload(c, oMF, 1, 3, S_old(c) + S_MF * N3))) return rc;
    if (c->mp.doMHD) {       // active cell B from the given face fluxes
        enq_bflux2fld(c, S_cur(c));
        enq_bflux2fld(c, S_old(c));
    }
    enq_bcs(c, S_cur(c));    // ghosts from the given interiors, oState's at its time
    enq_bcs(c, S_old(c), true, a->oclk);
    if (c->mp.doMHD) {       // cell B from the face fluxes
        enq_bflux2fld(c, S_cur(c));
        enq_bflux2fld(c, S_old(c));
    }
    if (mode == 1 && dts[0] == 0.0) {   // the first dt: CalcDT of the start, as gam_init_run
        launch_calc_dt(c->stream, c->g, c->mp, dt_box(c->g), c->clk, S_cur(c), c->met, c->bg);
        launch_finalize(c->stream, c->mp, c->clk, FIN_CALCDT | FIN_SETDT | FIN_INIT);
    }
    const int prof = c->prof;
    c->prof = 0;
    for (int n = 0; n < a->nsteps; n++) {
        if (mode == 1 && dts[n] > 0.0) launch_set_dt_val(c->stream, c->clk, a->dts, n);
        enq_step(c);
        if (mode == 0) launch_set_dt(c->stream, c->clk, a->clkh + n + 1);
    }
    c->prof = prof;
    HIPCHK(hipGetLastError());
    if ((rc = adj_objective(a, S_cur(c), a->lam + 2 * ADJ_SET * N3, J))) return rc;
    return adj_restore_final(a);
}

extern "C" int gam_adj_replay(gam_adj *a, const double *Gas, const double *oGas, const double *MF,
                              const double *oMF, double *J)
{
    if (!a || !Gas) return fail(GAM_E_ARG, "gam_adj_replay: NULL argument");
    if (!a->recorded) return fail(GAM_E_STATE, "gam_adj_replay: no recorded forward run");
    return adj_replay(a, Gas, oGas, MF, oMF, 0, nullptr, J);
}

extern "C" int gam_adj_replay_dt(gam_adj *a, const double *Gas, const double *oGas, const double *MF,
                                 const double *oMF, const double *dts, double *J)
{
    if (!a || !Gas) return fail(GAM_E_ARG, "gam_adj_replay_dt: NULL argument");
    if (!a->recorded) return fail(GAM_E_STATE, "gam_adj_replay_dt: no recorded forward run");
    if (dts)
        for (int n = 0; n < a->nsteps; n++)
            if (!(dts[n] >= 0.0)) return fail(GAM_E_ARG, "gam_adj_replay_dt: dts[%d] is negative or NaN", n);
    return adj_replay(a, Gas, oGas, MF, oMF, dts ? 1 : 2, dts, J);
}

extern "C" int gam_adj_set_dt_mode(gam_adj *a, int mode)
{
    if (!a) return fail(GAM_E_ARG, "gam_adj_set_dt_mode: NULL argument");
    if (mode < 0 || mode > 3) return fail(GAM_E_ARG, "gam_adj_set_dt_mode: mode must be 0 .. 3");
    if (mode && !a->pdt) {
        const size_t pl = (size_t)a->c->g.N3 * sizeof(double);
        hipError_t e = hipMalloc((void **)&a->pdt, 2 * pl);
        if (e == hipSuccess) e = hipMalloc((void **)&a->mu, (size_t)(a->nsteps + 1) * sizeof(double));
        if (e == hipSuccess) e = hipMalloc((void **)&a->dtpart, 2 * (size_t)dt_part_blocks(a->c->g) * sizeof(double));
        if (e == hipSuccess) e = hipMalloc((void **)&a->dtrec, (size_t)(a->nsteps + 1) * DT_REC * sizeof(double));
        if (e == hipSuccess)
            e = hipMalloc((void **)&a->dtrpart, 2 * (size_t)dt_record_blocks(a->c->g, dt_box(a->c->g)) * sizeof(double));
        if (e != hipSuccess) {
            for (double **b : {&a->pdt, &a->mu, &a->dtpart, &a->dtrec, &a->dtrpart}) {
                if (*b) hipFree(*b);
                *b = nullptr;
            }
            return fail(GAM_E_NOMEM, "gam_adj_set_dt_mode: %s", hipGetErrorString(e));
        }
    }
    a->dtMode = mode;   // modes 2 and 3 take the argmins the next forward records
    a->dtDone = 0;
    return GAM_OK;
}

extern "C" int gam_adj_step_dt(gam_adj *a, double *dts)
{
    if (!a || !dts) return fail(GAM_E_ARG, "gam_adj_step_dt: NULL argument");
    if (!a->recorded) return fail(GAM_E_STATE, "gam_adj_step_dt: no recorded forward run");
    std::vector<Clock> h(a->nsteps);
    HIPCHK(hipMemcpy(h.data(), a->clkh, (size_t)a->nsteps * sizeof(Clock), hipMemcpyDeviceToHost));
    for (int n = 0; n < a->nsteps; n++) dts[n] = h[n].dtA;
    return GAM_OK;
}

extern "C" int gam_adj_dt_grad(gam_adj *a, double *g)
{
    if (!a || !g) return fail(GAM_E_ARG, "gam_adj_dt_grad: NULL argument");
    if (!a->dtDone) return fail(GAM_E_STATE, "gam_adj_dt_grad: no backward sweep with the dt sensitivities on");
    HIPCHK(hipStreamSynchronize(a->c->stream));
    HIPCHK(hipMemcpy(g, a->mu, (size_t)a->nsteps * sizeof(double), hipMemcpyDeviceToHost));
    return GAM_OK;
}

extern "C" int gam_adj_wind_grad(gam_adj *a, double *gQ, double *gB)
{
    if (!a || !gQ || !gB) return fail(GAM_E_ARG, "gam_adj_wind_grad: NULL argument");
    if (!a->nw) return fail(GAM_E_STATE, "gam_adj_wind_grad: the model has no inflow BC");
    gam_ctx *c = a->c;
    HIPCHK(hipStreamSynchronize(c->stream));
    // the kernels add into WIND_COPIES copies of each array: sum them in order
    const size_t nq = (size_t)a->nw * 5, nb = (size_t)a->nw * 3;
    std::vector<double> hq(WIND_COPIES * nq), hb(WIND_COPIES * nb);
    HIPCHK(hipMemcpy(hq.data(), a->gwq, hq.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hb.data(), a->gwb, hb.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t e = 0; e < nq; e++) {
        double v = 0.0;
        for (int q = 0; q < WIND_COPIES; q++) v += hq[q * nq + e];
        gQ[e] = v;
    }
    for (size_t e = 0; e < nb; e++) {
        double v = 0.0;
        for (int q = 0; q < WIND_COPIES; q++) v += hb[q * nb + e];
        gB[e] = v;
    }
    return GAM_OK;
}

extern "C" int gam_adj_set_ostate_bc_time(gam_adj *a, int at_state_time)
{
    if (!a) return fail(GAM_E_ARG, "gam_adj_set_ostate_bc_time: NULL argument");
    a->ostAtState = at_state_time ? 1 : 0;
    return GAM_OK;
}

extern "C" int gam_adj_tape_info(gam_adj *a, int *taped_steps, int *overlap)
{
    if (!a) return fail(GAM_E_ARG, "gam_adj_tape_info: NULL argument");
    if (taped_steps) *taped_steps = a->itape ? a->Kt : 0;
    if (overlap) *overlap = a->itape2 ? 1 : 0;
    return GAM_OK;
}

extern "C" int gam_adj_times(gam_adj *a, double *fwd_ms, double *bwd_ms)
{
    if (!a) return fail(GAM_E_ARG, "gam_adj_times: NULL argument");
    if (fwd_ms) *fwd_ms = a->fwd_ms;
    if (bwd_ms) *bwd_ms = a->bwd_ms;
    return GAM_OK;
}
