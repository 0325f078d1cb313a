/*
 * gam_oracle.h -- CPU restatement of the GAMERA MHD time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path (libgamhip) is compared against.  Only tests/, the smoke()
 * entry point and bench.py's cpu_baseline leg may load it; the product
 * never links, loads or calls it.
 *
 * Every routine restates the reference Fortran (rgorby/MAGE, "Kaiju") loop
 * by loop, in the same operation order, compiled without fast-math and
 * without FMA contraction.  Citations are path:line into the reference.
 *
 * Memory layout is the reference's own (Fortran, column-major, i fastest):
 *   cell arrays  (isg:ieg, jsg:jeg, ksg:keg, comp)         -> nc per comp
 *   face arrays  (isg:ieg+1, jsg:jeg+1, ksg:keg+1, comp)   -> nf per comp
 * Multi-index components follow Fortran order too, e.g. Tf(i,j,k,c,d)
 * lives at [ijk + nf*(c + 9*d)].
 */
#ifndef GAM_ORACLE_H
#define GAM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NVAR 5

/* Boundary kinds (bcs.F90, wind.F90) */
enum {
    ORC_BC_PERIODIC = 0,   /* bcs.F90:279-450 */
    ORC_BC_ZEROGRAD = 1,   /* bcs.F90:799-930 */
    ORC_BC_SPHIN    = 2,   /* bcs.F90:934-957 (inner i only) */
    ORC_BC_SPHOUT   = 3,   /* bcs.F90:959-982 (outer i only) */
    ORC_BC_WIND     = 4    /* wind.F90:206-374 (outer i only) */
};

/* Recon methods (init.F90:442-454) */
enum { ORC_RECON_8CENT = 0, ORC_RECON_7UP = 1 };

/* CalcDT status flags (step.F90:103-156) */
enum { ORC_DT_OK = 0, ORC_DT_BAD = 1, ORC_DT_DISASTER = 2 };

typedef struct {
    int Nip, Njp, Nkp, nG;
    int Ni, Nj, Nk;
    int is, ie, js, je, ks, ke;
    int isg, ieg, jsg, jeg, ksg, keg;
    int isDT, ieDT, jsDT, jeDT, ksDT, keDT;
    long nc, nf;
} orc_dims;

typedef struct {
    double gamma, CFL, Vd0, Ca, cHogH, cHogM, pdmb, dFloor, pFloor;
    double t, tFin, dt, dt0, limDT0;
    int ts;
    int doMHD, doBoris, doBackground, doLFMLim, rmeth;
    int bc[6];                /* INI, OUTI, INJ, OUTJ, INK, OUTK */
    /* Inflow (wind) time series, discrete walls (wind.F90:766-868) */
    int nWind;
    const double *tW;         /* nWind */
    const double *Qw;         /* nWind x 5 primitives, column-major (nWind,5) */
    const double *Bw;         /* nWind x 3, column-major (nWind,3) */
    double Bx0, ByC, BzC;
    double xyzW[3];
} orc_model;

typedef struct {
    double *xyz;      /* nf*3   corners */
    double *xyzcc;    /* nc*3   */
    double *xfc;      /* nf*9   xfc(i,j,k,xyz,d) */
    double *di, *dj, *dk, *volume;   /* nc */
    double *Tf;       /* nf*27  */
    double *Te;       /* nf*27  */
    double *Teb;      /* nf*12  */
    double *face;     /* nf*3   */
    double *edge;     /* nf*3   */
    double *fcB0;     /* nc*9   (optional) */
    double *edgB0;    /* nc*6   */
    double *B0;       /* nc*3   */
    double *bFlux0;   /* nc*3   */
    double *dpB0;     /* nc*3   */
} orc_grid;

typedef struct {
    double *Gas;      /* nc*5 */
    double *magFlux;  /* nf*3 */
    double *Efld;     /* nc*3 */
    double *Bxyz;     /* nc*3 */
    double time;
} orc_state;

typedef struct orc_solver orc_solver;

void   orc_dims_init(orc_dims *D, int Nip, int Njp, int Nkp);

/* Cell-level kernels exposed for unit pinning */
void   orc_cell_c2p(const orc_model *M, const double *con, double *prim);
void   orc_cell_p2c(const orc_model *M, const double *prim, double *con);
double orc_pdm(double pdmb, double q0, double q1, double q2, double qI);
void   orc_pdm_lr(double pdmb, const double *nq4, double qi, double *ql, double *qr);
void   orc_recon_lr(const orc_model *M, const double *dV8, const double *nQ4,
                    const double *Q8, double *ql, double *qr);
void   orc_gas_kin_flux(const orc_model *M, int sgn, const double *prim,
                        const double *Tf9, double *gF5, double *dW5);
void   orc_mag_kin_flux(const orc_model *M, int sgn, const double *prim,
                        const double *mag3, double Bn, const double *Tf9,
                        const double *B03, double B0n, double *mF3,
                        double *bbD, double *VaD);
double orc_cell_dt(const orc_dims *D, const orc_model *M, const orc_grid *G,
                   const orc_state *S, int i, int j, int k);

/* Grid pre-compute */
void   orc_corners2grid(const orc_dims *D, orc_grid *G);
void   orc_add_dipole_b0(const orc_dims *D, orc_grid *G, double M0);
void   orc_dipole(double M0, double x, double y, double z, double *B);

/* Step pieces */
orc_solver *orc_solver_new(const orc_dims *D);
void   orc_solver_free(orc_solver *S);
void   orc_bflux2fld(const orc_dims *D, const orc_model *M, const orc_grid *G,
                     const double *magFlux, double *Bxyz);
int    orc_calc_dt(const orc_dims *D, orc_model *M, const orc_grid *G,
                   const orc_state *S, double *dt);
void   orc_enforce_bcs(const orc_dims *D, const orc_model *M, const orc_grid *G,
                       orc_state *S);
void   orc_advance_mhd(const orc_dims *D, orc_model *M, const orc_grid *G,
                       orc_state *S, orc_state *oS, orc_solver *W, double dt);
int    orc_step(const orc_dims *D, orc_model *M, const orc_grid *G,
                orc_state *S, orc_state *oS, orc_solver *W);
int    orc_init_run(const orc_dims *D, orc_model *M, const orc_grid *G,
                    orc_state *S, orc_state *oS);
double orc_divb(const orc_dims *D, const orc_model *M, const orc_grid *G,
                const orc_state *S, double *divout);
/* Static pieces of the step exposed for the oracle's pin tests */
void   orc_cell_boris(const orc_model *M, double *U2, const double *U0, const double *B2,
                      const double *B0, const double *dGasH, const double *dGasM, double dt);
void   orc_wind_interp(const orc_model *M, double t, double *out8);
void   orc_wind_front(const orc_model *M, double t, double *vFr);
void   orc_wind_at(const orc_model *M, const double *vFr, const double *xyz, double t, double *out8);
double orc_wind_wgt(const orc_model *M, const double *vFr, const double *xyz, double t);
double orc_ramp_down(double r, double rC, double lC);
/* Access to solver intermediates for staged comparisons */
double *orc_solver_array(orc_solver *W, int which);
long   orc_nuke_count(int reset);
enum { ORC_W_GFLX = 0, ORC_W_MFLX = 1, ORC_W_VF = 2, ORC_W_DGASH = 3,
       ORC_W_DGASM = 4, ORC_W_HF_GAS = 5, ORC_W_HF_MAGFLUX = 6,
       ORC_W_HF_BXYZ = 7 };

#ifdef __cplusplus
}
#endif
#endif
