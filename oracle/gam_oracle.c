/*
 * gam_oracle.c -- CPU restatement of the GAMERA MHD time step (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Loaded
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never
 * by the product.
 *
 * Built with gcc -O2 -fno-fast-math -ffp-contract=off so that every
 * expression is evaluated in the order written, mirroring Fortran's
 * left-to-right evaluation of the reference.  Citations are path:line into
 * the reference tree (rgorby/MAGE).
 *
 * Literal conventions (documented choices):
 *  - Fortran default-real literals that meet double-precision operands are
 *    taken at double precision (the Intel "-fpconstant" reading,
 *    cmake/compilers.cmake:84); every such literal on the path is exact in
 *    single precision anyway.
 *  - sqrt(2.0) in the Boris "nuke" threshold (stress.F90:405) is an
 *    intrinsic of a default-real argument, so its result is default real
 *    (single precision) under both the GNU and the Intel flags of the
 *    reference (-fpconstant does not change an intrinsic's result kind):
 *    the divisor is the single-precision sqrt(2), 1.41421353816986083984375.
 *  - norm2() is evaluated as sqrt(x*x+y*y+z*z) (no scaling).
 *  - x**2.0 is evaluated as x*x (the correctly rounded square).
 */
#include "gam_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TINY 1.0e-12           /* kdefs.F90:20 */
#define HUGEV 1.0e+12          /* kdefs.F90:20 */
#define PI (4.0 * atan(1.0))   /* kdefs.F90:18 */
#define CLIM 1.5               /* stress.F90:25 */
#define SQRT2_SP 1.41421353816986083984375   /* sqrt(2.0), default real (stress.F90:405) */
#define MINMACH 1.0            /* wind.F90:18 */

enum { DEN = 0, MOMX = 1, MOMY = 2, MOMZ = 3, ENERGY = 4 };
enum { VELX = 1, VELY = 2, VELZ = 3, PRESSURE = 4 };
enum { IDIR = 0, JDIR = 1, KDIR = 2 };
enum { NORMX = 0, NORMY = 1, NORMZ = 2, TAN1X = 3, TAN1Y = 4, TAN1Z = 5,
       TAN2X = 6, TAN2Y = 7, TAN2Z = 8 };
enum { XNQI = 0, YNQI = 1, XNQJ = 2, YNQJ = 3 };

/* recon.F90:22-23, 29 : weights are integer/real, element by element */
static double W8[8], W6[8], UP7[7];
static int weights_ready = 0;
static void init_weights(void)
{
    if (weights_ready) return;
    const int w8[8] = {-3, 29, -139, 533, 533, -139, 29, -3};
    const int w6[8] = {0, 1, -8, 37, 37, -8, 1, 0};
    const int u7[7] = {-3, 25, -101, 319, 214, -38, 4};
    for (int n = 0; n < 8; n++) { W8[n] = (double)w8[n] / 840.0; W6[n] = (double)w6[n] / 60.0; }
    for (int n = 0; n < 7; n++) UP7[n] = (double)u7[n] / 420.0;
    weights_ready = 1;
}

/* 12-point Gauss-Legendre (quadrature.F90:9-26) */
static const double gA[12] = {
    0.1252334085114689154724, 0.3678314989981801937527,
    0.5873179542866174472967, 0.7699026741943046870369,
    0.9041172563704748566785, 0.9815606342467192506906,
   -0.1252334085114689154724,-0.3678314989981801937527,
   -0.5873179542866174472967,-0.7699026741943046870369,
   -0.9041172563704748566785,-0.9815606342467192506906 };
static const double gW[12] = {
    0.2491470458134027850006, 0.233492536538354808761,
    0.203167426723065921749,  0.160078328543346226335,
    0.1069393259953184309603, 0.0471753363865118271946,
    0.2491470458134027850006, 0.233492536538354808761,
    0.203167426723065921749,  0.160078328543346226335,
    0.1069393259953184309603, 0.0471753363865118271946 };

/* ---------------------------------------------------------------- index */
static inline long CI(const orc_dims *D, int i, int j, int k)
{
    return (long)(i - D->isg) + (long)D->Ni * ((long)(j - D->jsg) + (long)D->Nj * (long)(k - D->ksg));
}
static inline long FI(const orc_dims *D, int i, int j, int k)
{
    return (long)(i - D->isg) + (long)(D->Ni + 1) * ((long)(j - D->jsg) + (long)(D->Nj + 1) * (long)(k - D->ksg));
}
static const int EIJK[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

static inline double dot8(const double *w, const double *q)
{
    double s = 0.0;
    for (int n = 0; n < 8; n++) s = s + w[n] * q[n];
    return s;
}
static inline double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline double norm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static inline void cross3(const double *a, const double *b, double *c)
{   /* math.F90:265-274 */
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline void normvec(const double *a, double *o)
{   /* math.F90:332-343 */
    double nA = sqrt(dot3(a, a));
    if (nA <= TINY) { o[0] = o[1] = o[2] = 0.0; }
    else { o[0] = a[0] / nA; o[1] = a[1] / nA; o[2] = a[2] / nA; }
}
static inline double fsign1(double x) { return copysign(1.0, x); }
static inline double fmaxd(double a, double b) { return (a > b) ? a : b; }
static inline double fmind(double a, double b) { return (a < b) ? a : b; }

/* ------------------------------------------------------------- dims */
void orc_dims_init(orc_dims *D, int Nip, int Njp, int Nkp)
{   /* init.F90:641-656, SetDomain init.F90:511-530 (single tile) */
    init_weights();
    D->Nip = Nip; D->Njp = Njp; D->Nkp = Nkp; D->nG = 4;
    D->Ni = Nip + 8; D->Nj = Njp + 8; D->Nk = Nkp + 8;
    D->is = 1; D->ie = Nip; D->js = 1; D->je = Njp; D->ks = 1; D->ke = Nkp;
    D->isg = D->is - 4; D->ieg = D->ie + 4;
    D->jsg = D->js - 4; D->jeg = D->je + 4;
    D->ksg = D->ks - 4; D->keg = D->ke + 4;
    D->isDT = D->is - 1; D->ieDT = D->ie + 1;
    D->jsDT = D->js - 1; D->jeDT = D->je + 1;
    D->ksDT = D->ks - 1; D->keDT = D->ke + 1;
    D->nc = (long)D->Ni * D->Nj * D->Nk;
    D->nf = (long)(D->Ni + 1) * (D->Nj + 1) * (D->Nk + 1);
}

/* ------------------------------------------------- cell conversions */
void orc_cell_c2p(const orc_model *M, const double *con, double *prim)
{   /* gamutils.F90:301-321 CellC2P: divides by unfloored rho */
    double rho = con[DEN];
    double Vx = con[MOMX] / rho, Vy = con[MOMY] / rho, Vz = con[MOMZ] / rho;
    double E = con[ENERGY];
    double KinE = 0.5 * rho * (Vx * Vx + Vy * Vy + Vz * Vz);
    prim[DEN] = fmaxd(rho, M->dFloor);
    prim[VELX] = Vx; prim[VELY] = Vy; prim[VELZ] = Vz;
    prim[PRESSURE] = fmaxd((M->gamma - 1) * (E - KinE), M->pFloor);
}

void orc_cell_p2c(const orc_model *M, const double *prim, double *con)
{   /* gamutils.F90:362-383 CellP2C */
    double rho = fmaxd(prim[DEN], M->dFloor);
    double Vx = prim[VELX], Vy = prim[VELY], Vz = prim[VELZ];
    double IntE = fmaxd(prim[PRESSURE], M->pFloor) / (M->gamma - 1);
    double KinE = 0.5 * rho * (Vx * Vx + Vy * Vy + Vz * Vz);
    con[DEN] = rho;
    con[MOMX] = rho * Vx; con[MOMY] = rho * Vy; con[MOMZ] = rho * Vz;
    con[ENERGY] = IntE + KinE;
}

/* recon.F90:75-90 BlockCon2Prim: floors rho before dividing */
static inline void block_con2prim(const orc_model *M, const double *c, double *p)
{
    double D = fmaxd(c[DEN], M->dFloor);
    double Mx = c[MOMX], My = c[MOMY], Mz = c[MOMZ], E = c[ENERGY];
    double KinE = 0.5 * (Mx * Mx + My * My + Mz * Mz) / D;
    double P = fmaxd((M->gamma - 1) * (E - KinE), M->pFloor);
    p[DEN] = D; p[VELX] = Mx / D; p[VELY] = My / D; p[VELZ] = Mz / D; p[PRESSURE] = P;
}

/* ------------------------------------------------------ limiter */
double orc_pdm(double pdmb, double q0, double q1, double q2, double qI)
{   /* recon.F90:275-301 PDM */
    double maxQ = fmaxd(q1, q2), minQ = fmind(q1, q2);
    double qN = fmaxd(minQ, fmind(qI, maxQ));
    double dq0 = pdmb * (q1 - q0), dq1 = pdmb * (q2 - q1);
    double s0 = fsign1(dq0), s1 = fsign1(dq1);
    double s01 = fabs(s0 + s1);
    double dqL = qN - q1;
    return qN - s1 * fmaxd(0.0, fabs(dqL) - s01 * fabs(dq0));
}

void orc_pdm_lr(double pdmb, const double *nq, double qi, double *ql, double *qr)
{   /* recon.F90:231-272 pdmLR */
    double v0 = nq[0], v1 = nq[1], v2 = nq[2], v3 = nq[3];
    double maxV = fmaxd(v1, v2), minV = fmind(v1, v2);
    double vN = fmaxd(minV, fmind(qi, maxV));
    double dv0 = pdmb * (v1 - v0), dv1 = pdmb * (v2 - v1), dv2 = pdmb * (v3 - v2);
    double s0 = fsign1(dv0), s1 = fsign1(dv1), s2 = fsign1(dv2);
    double q0 = fabs(s0 + s1), q1 = fabs(s1 + s2);
    double dvL = vN - v1, dvR = v2 - vN;
    *ql = vN - s1 * fmaxd(0.0, fabs(dvL) - q0 * fabs(dv0));
    *qr = vN + s1 * fmaxd(0.0, fabs(dvR) - q1 * fabs(dv2));
}

void orc_recon_lr(const orc_model *M, const double *dV, const double *nQ,
                  const double *Q, double *ql, double *qr)
{   /* recon.F90:96-122 BlockLRs -> Cen8LRs 125-158 / Up7LRs 162-198 */
    init_weights();
    double Vi = dot8(W8, dV);
    double QdV[8];
    for (int n = 0; n < 8; n++) QdV[n] = dV[n] * Q[n];
    if (M->rmeth == ORC_RECON_7UP) {
        double l = 0.0, r = 0.0;
        l = QdV[0] * UP7[0] + QdV[1] * UP7[1] + QdV[2] * UP7[2] + QdV[3] * UP7[3]
          + QdV[4] * UP7[4] + QdV[5] * UP7[5] + QdV[6] * UP7[6];
        r = QdV[7] * UP7[0] + QdV[6] * UP7[1] + QdV[5] * UP7[2] + QdV[4] * UP7[3]
          + QdV[3] * UP7[4] + QdV[2] * UP7[5] + QdV[1] * UP7[6];
        l = l / Vi; r = r / Vi;
        *ql = orc_pdm(M->pdmb, nQ[0], nQ[1], nQ[2], l);
        *qr = orc_pdm(M->pdmb, nQ[3], nQ[2], nQ[1], r);
    } else {
        double Qi = dot8(W8, QdV) / Vi;
        orc_pdm_lr(M->pdmb, nQ, Qi, ql, qr);
    }
}

/* ------------------------------------------------------ fluxes */
void orc_gas_kin_flux(const orc_model *M, int sgn, const double *W,
                      const double *T, double *gF, double *dW)
{   /* stress.F90:432-511 GasKinFlux (one side) */
    double D = W[DEN], P = W[PRESSURE], Vx = W[VELX], Vy = W[VELY], Vz = W[VELZ];
    double E = 0.5 * D * (Vx * Vx + Vy * Vy + Vz * Vz) + P / (M->gamma - 1);
    double lambda = D / (2 * P);
    double Vn = T[NORMX] * Vx + T[NORMY] * Vy + T[NORMZ] * Vz;
    double Vt1 = T[TAN1X] * Vx + T[TAN1Y] * Vy + T[TAN1Z] * Vz;
    double Vt2 = T[TAN2X] * Vx + T[TAN2Y] * Vy + T[TAN2Z] * Vz;
    double s = (double)sgn;
    double Vn0 = 0.5 * erfc(s * sqrt(lambda) * Vn);
    double Vn1 = Vn * Vn0 - 0.5 * s * exp(-(lambda * Vn * Vn)) / sqrt(PI * lambda);
    gF[DEN] = gF[DEN] + D * Vn1;
    gF[ENERGY] = gF[ENERGY] + (E + 0.5 * P) * Vn1 + 0.5 * P * Vn0 * Vn;
    double fMn = D * Vn * Vn1 + P * Vn0;
    double fMt1 = D * Vt1 * Vn1;
    double fMt2 = D * Vt2 * Vn1;
    double fMx = T[NORMX] * fMn + T[TAN1X] * fMt1 + T[TAN2X] * fMt2;
    double fMy = T[NORMY] * fMn + T[TAN1Y] * fMt1 + T[TAN2Y] * fMt2;
    double fMz = T[NORMZ] * fMn + T[TAN1Z] * fMt1 + T[TAN2Z] * fMt2;
    gF[MOMX] = gF[MOMX] + fMx;
    gF[MOMY] = gF[MOMY] + fMy;
    gF[MOMZ] = gF[MOMZ] + fMz;
    dW[DEN] = dW[DEN] + s * D;
    dW[MOMX] = dW[MOMX] + s * D * Vx;
    dW[MOMY] = dW[MOMY] + s * D * Vy;
    dW[MOMZ] = dW[MOMZ] + s * D * Vz;
    dW[ENERGY] = dW[ENERGY] + s * E;
}

void orc_mag_kin_flux(const orc_model *M, int sgn, const double *W,
                      const double *Bm, double Bn, const double *T,
                      const double *B0, double B0n, double *mF,
                      double *bbD, double *VaD)
{   /* stress.F90:513-599 MagKinFlux (one side) */
    double D = W[DEN], P = W[PRESSURE];
    double Bx = Bm[0], By = Bm[1], Bz = Bm[2];
    double B0x = B0[0], B0y = B0[1], B0z = B0[2];
    double dPb = 0.5 * (Bx * Bx + By * By + Bz * Bz);
    double Va2, Vac2;
    if (M->doBackground) {
        double tx = Bx + B0x, ty = By + B0y, tz = Bz + B0z;
        Va2 = (tx * tx + ty * ty + tz * tz) / D;
        *bbD = *bbD + 0.5 * (tx * tx + ty * ty + tz * tz);
    } else {
        Va2 = 2 * dPb / D;
        *bbD = *bbD + 0.5 * (Bx * Bx + By * By + Bz * Bz);
    }
    if (M->doBoris) Vac2 = Va2 / (1 + (Va2 / (M->Ca * M->Ca)));
    else Vac2 = Va2;
    double Vx = W[VELX], Vy = W[VELY], Vz = W[VELZ];
    double Nx = T[NORMX], Ny = T[NORMY], Nz = T[NORMZ];
    double Vn = Nx * Vx + Ny * Vy + Nz * Vz;
    double lambda = 1 / (2 * P / D + Vac2);
    double s = (double)sgn;
    double Vn0 = 0.5 * erfc(s * sqrt(lambda) * Vn);
    *VaD = *VaD + 0.5 * sqrt(Vac2);
    mF[0] = mF[0] + Vn0 * (dPb * Nx - Bx * Bn);
    mF[1] = mF[1] + Vn0 * (dPb * Ny - By * Bn);
    mF[2] = mF[2] + Vn0 * (dPb * Nz - Bz * Bn);
    if (M->doBackground) {
        double BdB0 = B0x * Bx + B0y * By + B0z * Bz;
        mF[0] = mF[0] + Vn0 * (-(B0x * Bn) - Bx * B0n + BdB0 * Nx);
        mF[1] = mF[1] + Vn0 * (-(B0y * Bn) - By * B0n + BdB0 * Ny);
        mF[2] = mF[2] + Vn0 * (-(B0z * Bn) - Bz * B0n + BdB0 * Nz);
    }
}

/* ------------------------------------------------------ solver */
struct orc_solver {
    double *gFlx;   /* nc*5*3  gFlx(i,j,k,nv,d) */
    double *mFlx;   /* nc*3*3  mFlx(i,j,k,x,d) */
    double *Vf;     /* nc*3 */
    double *dGasH;  /* nc*5 */
    double *dGasM;  /* nc*3 */
    orc_state hf;   /* Solver%StateHf */
};

orc_solver *orc_solver_new(const orc_dims *D)
{   /* init.F90:486-508 initSolver */
    orc_solver *W = (orc_solver *)calloc(1, sizeof(orc_solver));
    W->gFlx = (double *)calloc((size_t)D->nc * 15, sizeof(double));
    W->mFlx = (double *)calloc((size_t)D->nc * 9, sizeof(double));
    W->Vf = (double *)calloc((size_t)D->nc * 3, sizeof(double));
    W->dGasH = (double *)calloc((size_t)D->nc * 5, sizeof(double));
    W->dGasM = (double *)calloc((size_t)D->nc * 3, sizeof(double));
    W->hf.Gas = (double *)calloc((size_t)D->nc * 5, sizeof(double));
    W->hf.magFlux = (double *)calloc((size_t)D->nf * 3, sizeof(double));
    W->hf.Efld = (double *)calloc((size_t)D->nc * 3, sizeof(double));
    W->hf.Bxyz = (double *)calloc((size_t)D->nc * 3, sizeof(double));
    return W;
}

void orc_solver_free(orc_solver *W)
{
    if (!W) return;
    free(W->gFlx); free(W->mFlx); free(W->Vf); free(W->dGasH); free(W->dGasM);
    free(W->hf.Gas); free(W->hf.magFlux); free(W->hf.Efld); free(W->hf.Bxyz);
    free(W);
}

double *orc_solver_array(orc_solver *W, int which)
{
    switch (which) {
    case ORC_W_GFLX: return W->gFlx;
    case ORC_W_MFLX: return W->mFlx;
    case ORC_W_VF: return W->Vf;
    case ORC_W_DGASH: return W->dGasH;
    case ORC_W_DGASM: return W->dGasM;
    case ORC_W_HF_GAS: return W->hf.Gas;
    case ORC_W_HF_MAGFLUX: return W->hf.magFlux;
    case ORC_W_HF_BXYZ: return W->hf.Bxyz;
    }
    return 0;
}

/* -------------------------------------------------- B from fluxes */
static void cell_bxyz(const orc_dims *D, const orc_grid *G, const double *mf,
                      int i, int j, int k, double *B)
{   /* gridutils.F90:648-680 gamCellBxyz */
    long nf = D->nf;
    double dV = G->volume[CI(D, i, j, k)];
    double fI = mf[FI(D, i, j, k) + 0 * nf];
    double fJ = mf[FI(D, i, j, k) + 1 * nf];
    double fK = mf[FI(D, i, j, k) + 2 * nf];
    double fIp = mf[FI(D, i + 1, j, k) + 0 * nf];
    double fJp = mf[FI(D, i, j + 1, k) + 1 * nf];
    double fKp = mf[FI(D, i, j, k + 1) + 2 * nf];
    double Div = fIp - fI + fJp - fJ + fKp - fK;
    for (int x = 0; x < 3; x++) {
        double xfI = G->xfc[FI(D, i, j, k) + nf * (x + 3 * IDIR)];
        double xfJ = G->xfc[FI(D, i, j, k) + nf * (x + 3 * JDIR)];
        double xfK = G->xfc[FI(D, i, j, k) + nf * (x + 3 * KDIR)];
        double xfIp = G->xfc[FI(D, i + 1, j, k) + nf * (x + 3 * IDIR)];
        double xfJp = G->xfc[FI(D, i, j + 1, k) + nf * (x + 3 * JDIR)];
        double xfKp = G->xfc[FI(D, i, j, k + 1) + nf * (x + 3 * KDIR)];
        double b = (fIp * xfIp + fJp * xfJp + fKp * xfKp - fI * xfI - fJ * xfJ - fK * xfK) / dV;
        b = b - Div * G->xyzcc[CI(D, i, j, k) + D->nc * x] / dV;
        B[x] = b;
    }
}

void orc_bflux2fld(const orc_dims *D, const orc_model *M, const orc_grid *G,
                   const double *magFlux, double *Bxyz)
{   /* gridutils.F90:281-301 bFlux2Fld */
    (void)M;
#pragma omp parallel for collapse(2)
    for (int k = D->ksg; k <= D->keg; k++)
        for (int j = D->jsg; j <= D->jeg; j++)
            for (int i = D->isg; i <= D->ieg; i++) {
                double B[3];
                cell_bxyz(D, G, magFlux, i, j, k, B);
                long c = CI(D, i, j, k);
                Bxyz[c] = B[0]; Bxyz[c + D->nc] = B[1]; Bxyz[c + 2 * D->nc] = B[2];
            }
}

/* ------------------------------------------------------- dt */
double orc_cell_dt(const orc_dims *D, const orc_model *M, const orc_grid *G,
                   const orc_state *S, int i, int j, int k)
{   /* gridutils.F90:49-106 CellDT */
    long c = CI(D, i, j, k), nc = D->nc;
    double rho = S->Gas[c + nc * DEN];
    double Vx = S->Gas[c + nc * MOMX] / rho;
    double Vy = S->Gas[c + nc * MOMY] / rho;
    double Vz = S->Gas[c + nc * MOMZ] / rho;
    double MagV = sqrt(Vx * Vx + Vy * Vy + Vz * Vz);
    double ke = 0.5 * rho * (MagV * MagV);
    double e = S->Gas[c + nc * ENERGY] - ke;
    double P = (M->gamma - 1) * e;
    double Cs = sqrt(M->gamma * P / rho);
    double Vfl = MagV, Valf = 0.0, Vdiff = 0.0;
    if (M->doMHD) {
        double Bx = S->Bxyz[c], By = S->Bxyz[c + nc], Bz = S->Bxyz[c + 2 * nc];
        if (M->doBackground) {
            Bx = Bx + G->B0[c]; By = By + G->B0[c + nc]; Bz = Bz + G->B0[c + 2 * nc];
        }
        double MagB = sqrt(Bx * Bx + By * By + Bz * Bz);
        Valf = MagB / sqrt(rho);
        if (M->doBoris) Valf = M->Ca * Valf / sqrt(M->Ca * M->Ca + Valf * Valf);
    }
    double vCFL = Vfl + sqrt(Cs * Cs + Valf * Valf) + Vdiff;
    double dl = fmind(fmind(G->di[c], G->dj[c]), G->dk[c]);
    return M->CFL * dl / vCFL;
}

int orc_calc_dt(const orc_dims *D, orc_model *M, const orc_grid *G,
                const orc_state *S, double *dtOut)
{   /* step.F90:58-168 CalcDT; failure returns a code instead of stop */
    double dtOld = M->dt, dtMin = HUGEV;
    int status = ORC_DT_OK;
#pragma omp parallel for collapse(2) reduction(min : dtMin)
    for (int k = D->ksDT; k <= D->keDT; k++)
        for (int j = D->jsDT; j <= D->jeDT; j++)
            for (int i = D->isDT; i <= D->ieDT; i++) {
                double d = orc_cell_dt(D, M, G, S, i, j, k);
                dtMin = fmind(d, dtMin);
            }
    if (dtMin == HUGEV) status = ORC_DT_DISASTER;
    double dt = dtMin;
    if (M->ts > 0) {
        if (dtOld / dtMin >= 10) status = status > ORC_DT_BAD ? status : ORC_DT_BAD;
        if ((M->dt0 > TINY) && (dtMin / M->dt0 <= M->limDT0)) status = status > ORC_DT_BAD ? status : ORC_DT_BAD;
        if (dtMin <= TINY) status = ORC_DT_DISASTER;
    }
    if ((M->t + dt) > M->tFin) dt = fmaxd(M->tFin - M->t, TINY);
    if (M->dt0 > TINY && dt > (10.0 * M->dt0)) dt = 10.0 * M->dt0;
    *dtOut = dt;
    return status;
}

/* ------------------------------------------------------ BCs */
static void copy_cell(const orc_dims *D, double *A, int ncomp, long dst, long src)
{
    for (int c = 0; c < ncomp; c++) A[dst + D->nc * c] = A[src + D->nc * c];
}

static void bc_periodic(const orc_dims *D, orc_state *S, int side)
{   /* bcs.F90:279-450 */
    long nc = D->nc, nf = D->nf;
    double *G5 = S->Gas, *B3 = S->Bxyz, *mf = S->magFlux;
    int Ng = D->nG;
#define ISCC(i, j, k) ((k) <= D->keg && (j) <= D->jeg && (i) <= D->ieg && (k) >= D->ksg && (j) >= D->jsg && (i) >= D->isg)
    switch (side) {
    case 0: /* periodic_ibcI */
        for (int k = D->ksg; k <= D->keg + 1; k++)
            for (int j = D->jsg; j <= D->jeg + 1; j++)
                for (int n = 1; n <= Ng; n++) {
                    int ig = D->is - n, is = D->ie - n + 1;
                    if (ISCC(ig, j, k)) {
                        copy_cell(D, G5, 5, CI(D, ig, j, k), CI(D, is, j, k));
                        copy_cell(D, B3, 3, CI(D, ig, j, k), CI(D, is, j, k));
                    }
                    for (int d = 0; d < 3; d++) mf[FI(D, ig, j, k) + nf * d] = mf[FI(D, is, j, k) + nf * d];
                }
        break;
    case 1: /* periodic_obcI */
        for (int k = D->ksg; k <= D->keg + 1; k++)
            for (int j = D->jsg; j <= D->jeg + 1; j++)
                for (int n = 1; n <= Ng; n++) {
                    int ig = D->ie + n, is = D->is + n - 1;
                    if (ISCC(ig, j, k)) {
                        copy_cell(D, G5, 5, CI(D, ig, j, k), CI(D, is, j, k));
                        copy_cell(D, B3, 3, CI(D, ig, j, k), CI(D, is, j, k));
                    }
                    mf[FI(D, ig, j, k) + nf * 1] = mf[FI(D, is, j, k) + nf * 1];
                    mf[FI(D, ig, j, k) + nf * 2] = mf[FI(D, is, j, k) + nf * 2];
                    mf[FI(D, ig + 1, j, k) + nf * 0] = mf[FI(D, D->is + n, j, k) + nf * 0];
                }
        break;
    case 2: /* periodic_ibcJ */
        for (int k = D->ksg; k <= D->keg + 1; k++)
            for (int i = D->isg; i <= D->ieg + 1; i++)
                for (int n = 1; n <= Ng; n++) {
                    int jg = D->js - n, js = D->je - n + 1;
                    if (ISCC(i, jg, k)) {
                        copy_cell(D, G5, 5, CI(D, i, jg, k), CI(D, i, js, k));
                        copy_cell(D, B3, 3, CI(D, i, jg, k), CI(D, i, js, k));
                    }
                    for (int d = 0; d < 3; d++) mf[FI(D, i, jg, k) + nf * d] = mf[FI(D, i, js, k) + nf * d];
                }
        break;
    case 3: /* periodic_obcJ */
        for (int k = D->ksg; k <= D->keg + 1; k++)
            for (int i = D->isg; i <= D->ieg + 1; i++)
                for (int n = 1; n <= Ng; n++) {
                    int jg = D->je + n, js = D->js + n - 1;
                    if (ISCC(i, jg, k)) {
                        copy_cell(D, G5, 5, CI(D, i, jg, k), CI(D, i, js, k));
                        copy_cell(D, B3, 3, CI(D, i, jg, k), CI(D, i, js, k));
                    }
                    mf[FI(D, i, jg, k) + nf * 0] = mf[FI(D, i, js, k) + nf * 0];
                    mf[FI(D, i, jg, k) + nf * 2] = mf[FI(D, i, js, k) + nf * 2];
                    mf[FI(D, i, jg + 1, k) + nf * 1] = mf[FI(D, i, D->js + n, k) + nf * 1];
                }
        break;
    case 4: /* periodic_ibcK */
        for (int j = D->jsg; j <= D->jeg + 1; j++)
            for (int i = D->isg; i <= D->ieg + 1; i++)
                for (int n = 1; n <= Ng; n++) {
                    int kg = D->ks - n, ks = D->ke - n + 1;
                    if (ISCC(i, j, kg)) {
                        copy_cell(D, G5, 5, CI(D, i, j, kg), CI(D, i, j, ks));
                        copy_cell(D, B3, 3, CI(D, i, j, kg), CI(D, i, j, ks));
                    }
                    for (int d = 0; d < 3; d++) mf[FI(D, i, j, kg) + nf * d] = mf[FI(D, i, j, ks) + nf * d];
                }
        break;
    case 5: /* periodic_obcK */
        for (int j = D->jsg; j <= D->jeg + 1; j++)
            for (int i = D->isg; i <= D->ieg + 1; i++)
                for (int n = 1; n <= Ng; n++) {
                    int kg = D->ke + n, ks = D->ks + n - 1;
                    if (ISCC(i, j, kg)) {
                        copy_cell(D, G5, 5, CI(D, i, j, kg), CI(D, i, j, ks));
                        copy_cell(D, B3, 3, CI(D, i, j, kg), CI(D, i, j, ks));
                    }
                    mf[FI(D, i, j, kg) + nf * 0] = mf[FI(D, i, j, ks) + nf * 0];
                    mf[FI(D, i, j, kg) + nf * 1] = mf[FI(D, i, j, ks) + nf * 1];
                    mf[FI(D, i, j, kg + 1) + nf * 2] = mf[FI(D, i, j, D->ks + n) + nf * 2];
                }
        break;
    }
#undef ISCC
    (void)nc;
}

static void bc_zerograd(const orc_dims *D, const orc_grid *G, orc_state *S, int side)
{   /* bcs.F90:799-930 */
    long nf = D->nf;
    double *G5 = S->Gas, *mf = S->magFlux;
    int Ng = D->nG;
    switch (side) {
    case 0: /* zeroGrad_ibcI */
        for (int k = D->ksg; k <= D->keg; k++)
            for (int j = D->jsg; j <= D->jeg; j++)
                for (int n = 1; n <= Ng; n++) {
                    copy_cell(D, G5, 5, CI(D, D->is - n, j, k), CI(D, D->is, j, k));
                    for (int d = 0; d < 3; d++) mf[FI(D, D->is - n, j, k) + nf * d] = mf[FI(D, D->is, j, k) + nf * d];
                }
        break;
    case 1: /* zeroGrad_obcI */
        for (int k = D->ksg; k <= D->keg; k++)
            for (int j = D->jsg; j <= D->jeg; j++) {
                double B[3];
                cell_bxyz(D, G, mf, D->ie, j, k, B);
                for (int n = 1; n <= Ng; n++) {
                    int ig = D->ie + n;
                    copy_cell(D, G5, 5, CI(D, ig, j, k), CI(D, D->ie, j, k));
                    int ii[3] = {ig + 1, ig, ig};
                    for (int d = 0; d < 3; d++) {
                        long f = FI(D, ii[d], j, k);
                        double nrm[3] = {G->Tf[f + nf * (NORMX + 9 * d)], G->Tf[f + nf * (NORMY + 9 * d)], G->Tf[f + nf * (NORMZ + 9 * d)]};
                        mf[f + nf * d] = G->face[f + nf * d] * dot3(B, nrm);
                    }
                }
            }
        break;
    case 2: /* zeroGrad_ibcJ */
        for (int k = D->ksg; k <= D->keg; k++)
            for (int n = 1; n <= Ng; n++)
                for (int i = D->isg; i <= D->ieg; i++) {
                    copy_cell(D, G5, 5, CI(D, i, D->js - n, k), CI(D, i, D->js, k));
                    for (int d = 0; d < 3; d++) mf[FI(D, i, D->js - n, k) + nf * d] = mf[FI(D, i, D->js, k) + nf * d];
                }
        break;
    case 3: /* zeroGrad_obcJ */
        for (int k = D->ksg; k <= D->keg; k++)
            for (int n = 1; n <= Ng; n++)
                for (int i = D->isg; i <= D->ieg; i++) {
                    copy_cell(D, G5, 5, CI(D, i, D->je + n, k), CI(D, i, D->je, k));
                    mf[FI(D, i, D->je + n, k) + nf * 0] = mf[FI(D, i, D->je, k) + nf * 0];
                    mf[FI(D, i, D->je + n, k) + nf * 2] = mf[FI(D, i, D->je, k) + nf * 2];
                    mf[FI(D, i, D->je + n + 1, k) + nf * 1] = mf[FI(D, i, D->je + 1, k) + nf * 1];
                }
        break;
    case 4: /* zeroGrad_ibcK */
        for (int n = 1; n <= Ng; n++)
            for (int j = D->jsg; j <= D->jeg; j++)
                for (int i = D->isg; i <= D->ieg; i++) {
                    copy_cell(D, G5, 5, CI(D, i, j, D->ks - n), CI(D, i, j, D->ks));
                    for (int d = 0; d < 3; d++) mf[FI(D, i, j, D->ks - n) + nf * d] = mf[FI(D, i, j, D->ks) + nf * d];
                }
        break;
    case 5: /* zeroGrad_obcK -- the k+1 face takes the J component (bcs.F90:925) */
        for (int n = 1; n <= Ng; n++)
            for (int j = D->jsg; j <= D->jeg; j++)
                for (int i = D->isg; i <= D->ieg; i++) {
                    copy_cell(D, G5, 5, CI(D, i, j, D->ke + n), CI(D, i, j, D->ke));
                    mf[FI(D, i, j, D->ke + n) + nf * 0] = mf[FI(D, i, j, D->ke) + nf * 0];
                    mf[FI(D, i, j, D->ke + n) + nf * 1] = mf[FI(D, i, j, D->ke) + nf * 1];
                    mf[FI(D, i, j, D->ke + n + 1) + nf * 2] = mf[FI(D, i, j, D->ke + 1) + nf * 1];
                }
        break;
    }
}

static void bc_sphin(const orc_dims *D, orc_state *S)
{   /* bcs.F90:934-957 SphIn */
    long nf = D->nf;
    for (int k = D->ksg; k <= D->keg; k++)
        for (int j = D->jsg; j <= D->jeg; j++)
            for (int n = 1; n <= D->nG; n++) {
                int ig = D->is - n;
                copy_cell(D, S->Gas, 5, CI(D, ig, j, k), CI(D, D->is, j, k));
                for (int d = 0; d < 3; d++) S->magFlux[FI(D, ig, j, k) + nf * d] = S->magFlux[FI(D, D->is, j, k) + nf * d];
                copy_cell(D, S->Bxyz, 3, CI(D, ig, j, k), CI(D, D->is, j, k));
            }
}

static void bc_sphout(const orc_dims *D, orc_state *S)
{   /* bcs.F90:959-982 SphOut */
    long nf = D->nf;
    for (int k = D->ksg; k <= D->keg; k++)
        for (int j = D->jsg; j <= D->jeg; j++)
            for (int n = 1; n <= D->nG; n++) {
                int ig = D->ie + n;
                copy_cell(D, S->Gas, 5, CI(D, ig, j, k), CI(D, D->ie, j, k));
                S->magFlux[FI(D, ig + 1, j, k) + nf * 0] = S->magFlux[FI(D, D->ie, j, k) + nf * 0];
                S->magFlux[FI(D, ig, j, k) + nf * 1] = S->magFlux[FI(D, D->ie, j, k) + nf * 1];
                S->magFlux[FI(D, ig, j, k) + nf * 2] = S->magFlux[FI(D, D->ie, j, k) + nf * 2];
                copy_cell(D, S->Bxyz, 3, CI(D, ig, j, k), CI(D, D->ie, j, k));
            }
}

/* ----- inflow (solar-wind analogue) wind.F90 --------------------------- */
static void wind_interp(const orc_model *M, double t, double *Rho, double *Pr, double *V, double *B)
{   /* wind.F90:766-809 InterpWind with discrete walls (doWindInterp=F) */
    int N = M->nWind;
    double tMin = M->tW[0], tMax = M->tW[0], dtW = HUGEV;
    for (int n = 0; n < N; n++) { tMin = fmind(tMin, M->tW[n]); tMax = fmaxd(tMax, M->tW[n]); }
    for (int n = 1; n < N; n++) dtW = fmind(dtW, M->tW[n] - M->tW[n - 1]);
    int i0;
    if (t >= tMax) i0 = N - 1;
    else if (t <= tMin) i0 = 0;
    else {
        i0 = (int)floor((t - tMin) / dtW);
        if (i0 + 1 > N - 1) i0 = N - 2;
    }
    /* w0 = 1, w1 = 0 (wind.F90:862-866) */
    *Rho = M->Qw[i0 + N * DEN];
    *Pr = M->Qw[i0 + N * PRESSURE];
    V[0] = M->Qw[i0 + N * VELX]; V[1] = M->Qw[i0 + N * VELY]; V[2] = M->Qw[i0 + N * VELZ];
    B[0] = M->Bw[i0]; B[1] = M->Bw[i0 + N]; B[2] = M->Bw[i0 + 2 * N];
    B[0] = M->Bx0 + M->ByC * B[1] + M->BzC * B[2];
    if (t <= tMin) { B[0] = B[1] = B[2] = 0.0; }
}

static void wind_front(const orc_model *M, double t, double *vFr)
{   /* wind.F90:114-155 RefreshWind */
    double D, P, V0[3], B[3];
    wind_interp(M, t, &D, &P, V0, B);
    double vSpd = -norm3(V0);
    double bcc = sqrt(M->ByC * M->ByC + M->BzC * M->BzC);
    double btt = sqrt(1 + bcc * bcc);
    if (bcc > TINY) {
        double CosBp = M->ByC / bcc, SinBp = M->BzC / bcc;
        double CosF = 1.0 / btt, SinF = bcc / btt;
        vFr[0] = vSpd * (CosF * CosF);
        vFr[1] = -vSpd * CosF * SinF * CosBp;
        vFr[2] = -vSpd * CosF * SinF * SinBp;
    } else {
        vFr[0] = vSpd; vFr[1] = 0.0; vFr[2] = 0.0;
    }
}

static void wind_at(const orc_model *M, const double *vFr, const double *xyz, double t,
                    double *Rho, double *Pr, double *V, double *B)
{   /* wind.F90:157-175 GetWindAt */
    double vMag = norm3(vFr);
    double dx[3] = {xyz[0] - M->xyzW[0], xyz[1] - M->xyzW[1], xyz[2] - M->xyzW[2]};
    double DelT = dot3(dx, vFr) / (vMag * vMag);
    wind_interp(M, t - DelT, Rho, Pr, V, B);
}

static double ramp_down(double r, double rC, double lC)
{   /* math.F90:97-111 RampDown */
    if (r <= rC) return 1.0;
    if (r >= rC + lC) return 0.0;
    double rScl = (r - rC) / lC;
    double eArg = -1.0 / (1.0 - rScl * rScl);
    return exp(eArg + 1.0);
}

static double wind_wgt(const orc_model *M, const double *vFr, const double *xyz, double t)
{   /* wind.F90:180-203 wgtWind */
    double nHat[3], vHat[3], Rho, P, V[3], B[3];
    normvec(xyz, nHat);
    wind_at(M, vFr, xyz, t, &Rho, &P, V, B);
    normvec(V, vHat);
    double mv[3] = {-vHat[0], -vHat[1], -vHat[2]};
    double dTh = acos(dot3(mv, nHat)) * 180.0 / PI;
    return ramp_down(dTh, 90.0, 90.0);
}

static void con2fast(const orc_model *M, const double *gCon, const double *B, double *Cx)
{   /* gamutils.F90:338-360 Con2Fast (+ CellPress2Cs 323-336) */
    double rho = fmaxd(gCon[DEN], M->dFloor);
    double Vx = gCon[MOMX] / rho, Vy = gCon[MOMY] / rho, Vz = gCon[MOMZ] / rho;
    double KinE = 0.5 * rho * (Vx * Vx + Vy * Vy + Vz * Vz);
    double P = fmaxd((M->gamma - 1) * (gCon[ENERGY] - KinE), M->pFloor);
    double Cs = sqrt(M->gamma * P / rho);
    double D = gCon[DEN];
    double MagB = norm3(B);
    double Va = MagB / sqrt(D);
    if (M->doBoris) Va = M->Ca * Va / sqrt(M->Ca * M->Ca + Va * Va);
    *Cx = sqrt(Cs * Cs + Va * Va);
}

static void bc_wind(const orc_dims *D, const orc_model *M, const orc_grid *G, orc_state *S)
{   /* wind.F90:206-374 WindBC */
    long nc = D->nc, nf = D->nf;
    double t = M->t, vFr[3];
    wind_front(M, t, vFr);
    int ip = D->ie;
    for (int k = D->ksg; k <= D->keg + 1; k++)
        for (int j = D->jsg; j <= D->jeg + 1; j++)
            for (int n = 1; n <= D->nG; n++) {
                int ig = D->ie + n;
                if (k <= D->keg && j <= D->jeg) {
                    long cg = CI(D, ig, j, k), cp = CI(D, ip, j, k);
                    double xcc[3] = {G->xyzcc[cg], G->xyzcc[cg + nc], G->xyzcc[cg + 2 * nc]};
                    double wSW = wind_wgt(M, vFr, xcc, t), wMHD = 1.0 - wSW;
                    double gWsw[5] = {0, 0, 0, 0, 0}, gWin[5] = {0, 0, 0, 0, 0};
                    if (wSW > TINY) {
                        double Dn, P, V[3], Bs[3];
                        wind_at(M, vFr, xcc, t, &Dn, &P, V, Bs);
                        gWsw[DEN] = Dn; gWsw[PRESSURE] = P;
                        gWsw[VELX] = V[0]; gWsw[VELY] = V[1]; gWsw[VELZ] = V[2];
                    }
                    if (wMHD > TINY) {
                        double gCon[5], Bin[3], Cx;
                        for (int v = 0; v < 5; v++) gCon[v] = S->Gas[cp + nc * v];
                        for (int x = 0; x < 3; x++) Bin[x] = S->Bxyz[cp + nc * x];
                        orc_cell_c2p(M, gCon, gWin);
                        con2fast(M, gCon, Bin, &Cx);
                        double V[3] = {gWin[VELX], gWin[VELY], gWin[VELZ]};
                        long fp = FI(D, ip + 1, j, k);
                        double nH[3] = {G->Tf[fp + nf * (NORMX)], G->Tf[fp + nf * (NORMY)], G->Tf[fp + nf * (NORMZ)]};
                        if (dot3(V, nH) < MINMACH * Cx) {
                            double vp = dot3(V, nH);
                            for (int x = 0; x < 3; x++) V[x] = V[x] - vp * nH[x] + MINMACH * Cx * nH[x];
                        }
                        gWin[VELX] = V[0]; gWin[VELY] = V[1]; gWin[VELZ] = V[2];
                    }
                    double gW[5], gCon[5];
                    for (int v = 0; v < 5; v++) gW[v] = wSW * gWsw[v] + wMHD * gWin[v];
                    orc_cell_p2c(M, gW, gCon);
                    for (int v = 0; v < 5; v++) S->Gas[cg + nc * v] = gCon[v];
                }
                /* Face fluxes: I at ig+1, J and K at ig */
                int fi[3] = {ig + 1, ig, ig};
                int ii[3] = {ip + 1, ip, ip};
                for (int d = 0; d < 3; d++) {
                    long f = FI(D, fi[d], j, k);
                    double xf[3] = {G->xfc[f + nf * (0 + 3 * d)], G->xfc[f + nf * (1 + 3 * d)], G->xfc[f + nf * (2 + 3 * d)]};
                    double wSW = wind_wgt(M, vFr, xf, t);
                    double Dn, P, V[3], Bs[3];
                    wind_at(M, vFr, xf, t, &Dn, &P, V, Bs);
                    double nH[3] = {G->Tf[f + nf * (NORMX + 9 * d)], G->Tf[f + nf * (NORMY + 9 * d)], G->Tf[f + nf * (NORMZ + 9 * d)]};
                    double swFlx = G->face[f + nf * d] * dot3(nH, Bs);
                    double inFlx = S->magFlux[FI(D, ii[d], j, k) + nf * d];
                    S->magFlux[f + nf * d] = wSW * swFlx + (1 - wSW) * inFlx;
                }
            }
    /* Ghost Bxyz from the blended fluxes, k mapped periodically (ringutils.F90:482-518) */
    for (int k = D->ksg; k <= D->keg; k++)
        for (int j = D->jsg; j <= D->jeg; j++)
            for (int n = 1; n <= D->nG; n++) {
                int ig = D->ie + n, kp = k;
                if (k < D->ks) kp = k + D->Nkp;
                else if (k > D->ke) kp = k - D->Nkp;
                double B[3];
                cell_bxyz(D, G, S->magFlux, ig, j, kp, B);
                long c = CI(D, ig, j, k);
                S->Bxyz[c] = B[0]; S->Bxyz[c + nc] = B[1]; S->Bxyz[c + 2 * nc] = B[2];
            }
}

void orc_enforce_bcs(const orc_dims *D, const orc_model *M, const orc_grid *G, orc_state *S)
{   /* step.F90:17-55 EnforceBCs: INI,OUTI,INJ,OUTJ,INK,OUTK in order */
    for (int side = 0; side < 6; side++) {
        switch (M->bc[side]) {
        case ORC_BC_PERIODIC: bc_periodic(D, S, side); break;
        case ORC_BC_ZEROGRAD: bc_zerograd(D, G, S, side); break;
        case ORC_BC_SPHIN: bc_sphin(D, S); break;
        case ORC_BC_SPHOUT: bc_sphout(D, S); break;
        case ORC_BC_WIND: bc_wind(D, M, G, S); break;
        }
    }
}

/* ------------------------------------------------------ predictor */
static void predictor(const orc_dims *D, const orc_model *M, const orc_grid *G,
                      const orc_state *S, const orc_state *oS, orc_state *pS, double pdt)
{   /* mhdgroup.F90:444-524 Predictor, 528-561 PredictCell, 709-717 ExtrapAB2 */
    double odt = S->time - oS->time;
    long nf = D->nf, nc = D->nc;
    if (M->doMHD) {
#pragma omp parallel for collapse(2)
        for (int k = D->ksg; k <= D->keg + 1; k++)
            for (int j = D->jsg; j <= D->jeg + 1; j++)
                for (int i = D->isg; i <= D->ieg + 1; i++) {
                    long f = FI(D, i, j, k);
                    for (int d = 0; d < 3; d++) {
                        double Q = S->magFlux[f + nf * d], oQ = oS->magFlux[f + nf * d];
                        double dQ = (Q - oQ) / odt;
                        pS->magFlux[f + nf * d] = Q + pdt * dQ;
                    }
                }
    }
    orc_bflux2fld(D, M, G, pS->magFlux, pS->Bxyz);
#pragma omp parallel for collapse(2)
    for (int k = D->ksg; k <= D->keg; k++)
        for (int j = D->jsg; j <= D->jeg; j++)
            for (int i = D->isg; i <= D->ieg; i++) {
                long c = CI(D, i, j, k);
                double Q[5], oQ[5], W[5], oW[5], pW[5], pQ[5];
                for (int v = 0; v < 5; v++) { Q[v] = S->Gas[c + nc * v]; oQ[v] = oS->Gas[c + nc * v]; }
                orc_cell_c2p(M, oQ, oW);
                orc_cell_c2p(M, Q, W);
                for (int v = 0; v < 5; v++) {
                    double dQ = (W[v] - oW[v]) / odt;
                    pW[v] = W[v] + pdt * dQ;
                }
                pW[DEN] = fmaxd(pW[DEN], M->dFloor);
                pW[PRESSURE] = fmaxd(pW[PRESSURE], M->pFloor);
                orc_cell_p2c(M, pW, pQ);
                for (int v = 0; v < 5; v++) pS->Gas[c + nc * v] = pQ[v];
            }
}

/* ------------------------------------------------------ E field */
static void mom2face(const orc_dims *D, const orc_model *M, const orc_grid *G,
                     const double *Gas, int i, int j, int k, int dT, double *Vf3)
{   /* fields.F90:235-285 Mom2Face */
    long nc = D->nc;
    double VolB[8], VelB[3][8];
    for (int n = 0; n < 8; n++) {
        int ii = i + (n - 4) * EIJK[dT][0], jj = j + (n - 4) * EIJK[dT][1], kk = k + (n - 4) * EIJK[dT][2];
        long c = CI(D, ii, jj, kk);
        double vol = G->volume[c], den = Gas[c + nc * DEN];
        VolB[n] = vol;
        for (int d = 0; d < 3; d++) VelB[d][n] = vol * Gas[c + nc * (MOMX + d)] / fmaxd(den, M->dFloor);
    }
    double dV = fmaxd(dot8(W8, VolB), TINY);
    for (int d = 0; d < 3; d++) Vf3[d] = dot8(W8, VelB[d]) / dV;
}

static void edge_lrs(const orc_dims *D, const orc_model *M, const orc_grid *G,
                     const double *bFn, const double *bFh, int i, int j, int k,
                     int dT1, int dT2, double *bL, double *bR, double *fA)
{   /* fields.F90:288-362 EdgeLRs (doBdA) */
    long nf = D->nf;
    double AreaB[8], MagB[8], nAreaB[4], nMagB[4], VolB[8];
    const double *bFlim = M->doLFMLim ? bFn : bFh;
    for (int n = 0; n < 8; n++) {
        int o = n - 4;
        long f = FI(D, i + o * EIJK[dT2][0], j + o * EIJK[dT2][1], k + o * EIJK[dT2][2]);
        AreaB[n] = G->face[f + nf * dT1];
        MagB[n] = bFh[f + nf * dT1];
    }
    for (int m = 0; m < 4; m++) {
        int o = m - 2;
        long f = FI(D, i + o * EIJK[dT2][0], j + o * EIJK[dT2][1], k + o * EIJK[dT2][2]);
        nAreaB[m] = G->face[f + nf * dT1];
        nMagB[m] = bFlim[f + nf * dT1];
    }
    for (int n = 0; n < 8; n++) { VolB[n] = AreaB[n]; MagB[n] = MagB[n] / fmaxd(AreaB[n], TINY); }
    for (int m = 0; m < 4; m++) nMagB[m] = nMagB[m] / fmaxd(nAreaB[m], TINY);
    double L, R;
    orc_recon_lr(M, VolB, nMagB, MagB, &L, &R);
    *fA = dot8(W8, AreaB);
    *bL = L * (*fA);
    *bR = R * (*fA);
}

static void calc_elec_field(const orc_dims *D, const orc_model *M, const orc_grid *G,
                            const orc_state *nS, const orc_state *hS, double *Vf, double *E)
{   /* fields.F90:70-232 CalcElecField */
    long nc = D->nc, nf = D->nf;
    memset(E, 0, sizeof(double) * (size_t)nc * 3);
    for (int eD = 0; eD < 3; eD++) {
        int dT1, dT2, ie, je, ke;
        if (eD == IDIR) { dT1 = JDIR; dT2 = KDIR; ie = D->ie; je = D->je + 1; ke = D->ke + 1; }
        else if (eD == JDIR) { dT1 = KDIR; dT2 = IDIR; ie = D->ie + 1; je = D->je; ke = D->ke + 1; }
        else { dT1 = IDIR; dT2 = JDIR; ie = D->ie + 1; je = D->je + 1; ke = D->ke; }
        /* Mom2Face sweep ranges (fields.F90:119-175) */
        int i0 = D->isg, i1 = D->ieg, j0 = D->jsg, j1 = D->jeg, k0 = D->ksg, k1 = D->keg;
        if (eD == IDIR) { j0 = D->js; j1 = D->je + 1; }
        else if (eD == JDIR) { k0 = D->ks; k1 = D->ke + 1; }
        else { i0 = D->is; i1 = D->ie + 1; }
#pragma omp parallel for collapse(2)
        for (int k = k0; k <= k1; k++)
            for (int j = j0; j <= j1; j++)
                for (int i = i0; i <= i1; i++) {
                    double v3[3];
                    mom2face(D, M, G, hS->Gas, i, j, k, dT1, v3);
                    long c = CI(D, i, j, k);
                    Vf[c] = v3[0]; Vf[c + nc] = v3[1]; Vf[c + 2 * nc] = v3[2];
                }
        /* VxB on active edges (fields.F90:187-224) */
        int e1[3] = {0, 0, 0}, e2[3] = {0, 0, 0};
        if (eD == IDIR) { e1[1] = 1; e2[2] = 1; }
        else if (eD == JDIR) { e1[0] = 1; e2[2] = 1; }
        else { e1[0] = 1; e2[1] = 1; }
#pragma omp parallel for collapse(2)
        for (int k = D->ks; k <= ke; k++)
            for (int j = D->js; j <= je; j++)
                for (int i = D->is; i <= ie; i++) {
                    long c = CI(D, i, j, k), f = FI(D, i, j, k);
                    /* GetCornerV (fields.F90:457-517) */
                    double AreaB[8], VelB[3][8], dA, Vc[3];
                    for (int n = 0; n < 8; n++) {
                        int o = n - 4;
                        int ii = i + o * EIJK[dT2][0], jj = j + o * EIJK[dT2][1], kk = k + o * EIJK[dT2][2];
                        AreaB[n] = G->face[FI(D, ii, jj, kk) + nf * dT1];
                        long cc = CI(D, ii, jj, kk);
                        for (int d = 0; d < 3; d++) VelB[d][n] = Vf[cc + nc * d];
                    }
                    dA = fmaxd(dot8(W8, AreaB), TINY);
                    for (int d = 0; d < 3; d++) {
                        double pr[8];
                        for (int n = 0; n < 8; n++) pr[n] = AreaB[n] * VelB[d][n];
                        Vc[d] = dot8(W8, pr) / dA;
                    }
                    const double *Te = G->Te;
                    double v1 = Vc[0] * Te[f + nf * (TAN1X + 9 * eD)] + Vc[1] * Te[f + nf * (TAN1Y + 9 * eD)] + Vc[2] * Te[f + nf * (TAN1Z + 9 * eD)];
                    double v2 = Vc[0] * Te[f + nf * (TAN2X + 9 * eD)] + Vc[1] * Te[f + nf * (TAN2Y + 9 * eD)] + Vc[2] * Te[f + nf * (TAN2Z + 9 * eD)];
                    double vDiff = sqrt(v1 * v1 + v2 * v2);
                    /* GetCornerD (fields.F90:427-454) */
                    const double *Dh = hS->Gas;
                    double Dc = 0.25 * (Dh[c] + Dh[CI(D, i - e1[0], j - e1[1], k - e1[2])]
                                        + Dh[CI(D, i - e2[0], j - e2[1], k - e2[2])]
                                        + Dh[CI(D, i - e1[0] - e2[0], j - e1[1] - e2[1], k - e1[2] - e2[2])]);
                    /* GetCornerB (fields.F90:365-424) */
                    double bT1L, bT1R, fA1, bT2L, bT2R, fA2;
                    edge_lrs(D, M, G, nS->magFlux, hS->magFlux, i, j, k, dT1, dT2, &bT1L, &bT1R, &fA1);
                    edge_lrs(D, M, G, nS->magFlux, hS->magFlux, i, j, k, dT2, dT1, &bT2L, &bT2R, &fA2);
                    double xni = G->Teb[f + nf * (XNQI + 4 * eD)], yni = G->Teb[f + nf * (YNQI + 4 * eD)];
                    double xnj = G->Teb[f + nf * (XNQJ + 4 * eD)], ynj = G->Teb[f + nf * (YNQJ + 4 * eD)];
                    double detT = 1.0 / (xni * ynj - xnj * yni);
                    double bT1 = 0.5 * (bT1L + bT1R) / fA1;
                    double bT2 = 0.5 * (bT2L + bT2R) / fA2;
                    double db1 = (bT1R - bT1L) / fA1;
                    double db2 = (bT2R - bT2L) / fA2;
                    double Jd = db2 - db1;
                    double b1 = detT * (ynj * bT1 - yni * bT2);
                    double b2 = detT * (-(xnj * bT1) + xni * bT2);
                    if (M->doBackground) {
                        b1 = b1 + G->edgB0[c + nc * (0 + 2 * eD)];
                        b2 = b2 + G->edgB0[c + nc * (1 + 2 * eD)];
                    }
                    double edge = G->edge[f + nf * eD];
                    double vA = sqrt(b1 * b1 + b2 * b2) / sqrt(Dc);
                    if (M->doBoris) vA = M->Ca * vA / sqrt(M->Ca * M->Ca + vA * vA);
                    vDiff = vDiff + vA;
                    vDiff = fmind(vDiff, M->CFL * edge / M->dt);
                    double e = -(v1 * b2 - v2 * b1) + M->Vd0 * vDiff * Jd;
                    E[c + nc * eD] = e * edge;
                }
    }
}

/* ------------------------------------------------------ stress */
/* Faces on which the nuke term fired since the last orc_nuke_count(1)
 * (instrumentation for tests: shows that a case exercises stress.F90:399-417). */
static long nuke_faces = 0;

long orc_nuke_count(int reset)
{
    long n = nuke_faces;
    if (reset) nuke_faces = 0;
    return n;
}

static void face_flux(const orc_dims *D, const orc_model *M, const orc_grid *G,
                      const orc_state *nS, const orc_state *hS, int i, int j, int k,
                      int dN, double *gFo, double *mFo)
{   /* stress.F90:204-429 Fluxes, one interface */
    long nc = D->nc, nf = D->nf;
    const int *e = EIJK[dN];
    double VolB[8], ConB[5][8], nConB[5][4], MagB[3][8], nMagB[3][4];
    const orc_state *lim = M->doLFMLim ? nS : hS;
    for (int n = 0; n < 8; n++) {
        int o = n - 4;
        long c = CI(D, i + o * e[0], j + o * e[1], k + o * e[2]);
        VolB[n] = G->volume[c];
        for (int v = 0; v < 5; v++) ConB[v][n] = hS->Gas[c + nc * v];
        for (int x = 0; x < 3; x++) MagB[x][n] = hS->Bxyz[c + nc * x];
    }
    for (int m = 0; m < 4; m++) {
        int o = m - 2;
        long c = CI(D, i + o * e[0], j + o * e[1], k + o * e[2]);
        for (int v = 0; v < 5; v++) nConB[v][m] = lim->Gas[c + nc * v];
        for (int x = 0; x < 3; x++) nMagB[x][m] = lim->Bxyz[c + nc * x];
    }
    long f = FI(D, i, j, k), cf = CI(D, i, j, k);
    double Tf[9];
    for (int q = 0; q < 9; q++) Tf[q] = G->Tf[f + nf * (q + 9 * dN)];
    double faB = G->face[f + nf * dN];

    double MagL[3] = {0, 0, 0}, MagR[3] = {0, 0, 0}, Bn = 0.0, B0[3] = {0, 0, 0}, B0n = 0.0;
    if (M->doMHD) {
        for (int x = 0; x < 3; x++) orc_recon_lr(M, VolB, nMagB[x], MagB[x], &MagL[x], &MagR[x]);
        Bn = hS->magFlux[f + nf * dN] / faB;
        if (M->doBackground) {
            for (int x = 0; x < 3; x++) B0[x] = G->fcB0[cf + nc * (x + 3 * dN)];
            B0n = G->bFlux0[cf + nc * dN] / faB;
        }
    }
    /* BlockStateLRs (recon.F90:39-93) */
    double PrimB[5][8], nPrimB[5][4], PrimL[5], PrimR[5];
    for (int n = 0; n < 8; n++) {
        double c5[5] = {ConB[0][n], ConB[1][n], ConB[2][n], ConB[3][n], ConB[4][n]}, p5[5];
        block_con2prim(M, c5, p5);
        for (int v = 0; v < 5; v++) PrimB[v][n] = p5[v];
    }
    for (int m = 0; m < 4; m++) {
        double c5[5] = {nConB[0][m], nConB[1][m], nConB[2][m], nConB[3][m], nConB[4][m]}, p5[5];
        block_con2prim(M, c5, p5);
        for (int v = 0; v < 5; v++) nPrimB[v][m] = p5[v];
    }
    for (int v = 0; v < 5; v++) orc_recon_lr(M, VolB, nPrimB[v], PrimB[v], &PrimL[v], &PrimR[v]);

    double mF[3] = {0, 0, 0}, bbD = 0.0, VaD = 0.0;
    if (M->doMHD) {
        orc_mag_kin_flux(M, -1, PrimL, MagL, Bn, Tf, B0, B0n, mF, &bbD, &VaD);
        orc_mag_kin_flux(M, +1, PrimR, MagR, Bn, Tf, B0, B0n, mF, &bbD, &VaD);
        if (M->doBoris) {
            for (int x = 0; x < 3; x++) {
                double deeV = PrimR[VELX + x] - PrimL[VELX + x];
                double deeP = bbD * deeV / (M->Ca * M->Ca);
                mF[x] = mF[x] - M->cHogM * VaD * deeP;
            }
        }
        for (int x = 0; x < 3; x++) mFo[x] = faB * mF[x];
    }
    double gF[5] = {0, 0, 0, 0, 0}, dW[5] = {0, 0, 0, 0, 0};
    orc_gas_kin_flux(M, -1, PrimL, Tf, gF, dW);
    orc_gas_kin_flux(M, +1, PrimR, Tf, gF, dW);
    for (int v = 0; v < 5; v++) gF[v] = gF[v] - 1.0 * M->cHogH * VaD * dW[v];
    if (M->doBoris) {  /* doNuke (stress.F90:399-417) */
        double vL = sqrt(PrimL[1] * PrimL[1] + PrimL[2] * PrimL[2] + PrimL[3] * PrimL[3]);
        double vR = sqrt(PrimR[1] * PrimR[1] + PrimR[2] * PrimR[2] + PrimR[3] * PrimR[3]);
        double Vn = sqrt(vL * vL + vR * vR) / SQRT2_SP;
        if (Vn >= CLIM * M->Ca) {
#pragma omp atomic
            nuke_faces++;
            for (int v = 0; v < 5; v++) {
                double dCon = ConB[v][4] - ConB[v][3];
                gF[v] = gF[v] - M->Ca * dCon;
            }
        }
    }
    for (int v = 0; v < 5; v++) gFo[v] = faB * gF[v];
}

static void calc_stress(const orc_dims *D, const orc_model *M, const orc_grid *G,
                        const orc_state *nS, const orc_state *hS, orc_solver *W)
{   /* stress.F90:41-198 CalcStress */
    long nc = D->nc;
    for (int d = 0; d < 3; d++) {
        int ie = D->ie + (d == IDIR), je = D->je + (d == JDIR), ke = D->ke + (d == KDIR);
#pragma omp parallel for collapse(2)
        for (int k = D->ks; k <= ke; k++)
            for (int j = D->js; j <= je; j++)
                for (int i = D->is; i <= ie; i++) {
                    double gF[5], mF[3] = {0, 0, 0};
                    face_flux(D, M, G, nS, hS, i, j, k, d, gF, mF);
                    long c = CI(D, i, j, k);
                    for (int v = 0; v < 5; v++) W->gFlx[c + nc * (v + 5 * d)] = gF[v];
                    if (M->doMHD) for (int x = 0; x < 3; x++) W->mFlx[c + nc * (x + 3 * d)] = mF[x];
                }
    }
    /* Flux2Deltas (stress.F90:152-194) */
#pragma omp parallel for collapse(2)
    for (int k = D->ks; k <= D->ke; k++)
        for (int j = D->js; j <= D->je; j++)
            for (int i = D->is; i <= D->ie; i++) {
                long c = CI(D, i, j, k);
                long cI = CI(D, i + 1, j, k), cJ = CI(D, i, j + 1, k), cK = CI(D, i, j, k + 1);
                double dV = G->volume[c];
                for (int v = 0; v < 5; v++) {
                    const double *g = W->gFlx;
                    W->dGasH[c + nc * v] = (g[c + nc * (v + 0)] - g[cI + nc * (v + 0)]
                                          + g[c + nc * (v + 5)] - g[cJ + nc * (v + 5)]
                                          + g[c + nc * (v + 10)] - g[cK + nc * (v + 10)]) / dV;
                }
                if (M->doMHD) {
                    for (int x = 0; x < 3; x++) {
                        const double *m = W->mFlx;
                        double r = (m[c + nc * (x + 0)] - m[cI + nc * (x + 0)]
                                  + m[c + nc * (x + 3)] - m[cJ + nc * (x + 3)]
                                  + m[c + nc * (x + 6)] - m[cK + nc * (x + 6)]) / dV;
                        if (M->doBackground) r = r + G->dpB0[c + nc * x];
                        W->dGasM[c + nc * x] = r;
                    }
                }
            }
}

/* ------------------------------------------------------ update */
static void cell_boris(const orc_model *M, double *U2, const double *U0,
                       const double *B2, const double *B0, const double *dGasH,
                       const double *dGasM, double Ca, double dt)
{   /* mhdgroup.F90:336-441 CellBoris (single fluid) */
    double W2[5];
    orc_cell_c2p(M, U2, W2);
    double D2 = fmaxd(W2[DEN], M->dFloor);
    double P2 = fmaxd(W2[PRESSURE], M->pFloor);
    double D0 = fmaxd(U0[DEN], M->dFloor);
    double V0[3] = {U0[MOMX] / D0, U0[MOMY] / D0, U0[MOMZ] / D0};
    double B1[3] = {0.5 * (B2[0] + B0[0]), 0.5 * (B2[1] + B0[1]), 0.5 * (B2[2] + B0[2])};
    double alpha = dot3(B1, B1) / (D2 * Ca * Ca);
    double dRho = dt * dGasH[DEN];
    double dPGas[3] = {dt * dGasH[MOMX], dt * dGasH[MOMY], dt * dGasH[MOMZ]};
    double dPMag[3] = {dt * dGasM[0], dt * dGasM[1], dt * dGasM[2]};
    double pCon = 1.0 / (1.0 + alpha);
    double bh1[3];
    normvec(B1, bh1);
    double dMom[3];
    if (alpha > TINY) {
        double sGM[3] = {dPGas[0] + dPMag[0], dPGas[1] + dPMag[1], dPGas[2] + dPMag[2]};
        double pG = dot3(dPGas, bh1), pS = dot3(sGM, bh1), pV = dot3(V0, bh1);
        for (int x = 0; x < 3; x++) {
            double para = pG * bh1[x];
            double perpS = sGM[x] - pS * bh1[x];
            double perpV = V0[x] - pV * bh1[x];
            dMom[x] = para + pCon * perpS + alpha * pCon * dRho * perpV;
        }
    } else {
        for (int x = 0; x < 3; x++) dMom[x] = dPGas[x] + dPMag[x];
    }
    for (int x = 0; x < 3; x++) W2[VELX + x] = (D0 * V0[x] + dMom[x]) / D2;
    W2[DEN] = D2;
    W2[PRESSURE] = P2;
    orc_cell_p2c(M, W2, U2);
}

static void cell_maxwell(const orc_model *M, double *U, const double *dGasM, double dt)
{   /* mhdgroup.F90:299-329 CellMaxwell */
    double D = fmaxd(U[DEN], M->dFloor);
    double Mx = U[MOMX], My = U[MOMY], Mz = U[MOMZ], TotE = U[ENERGY];
    double KinE = 0.5 * (Mx * Mx + My * My + Mz * Mz) / D;
    double P = fmaxd((M->gamma - 1) * (TotE - KinE), M->pFloor);
    Mx = Mx + dt * dGasM[0];
    My = My + dt * dGasM[1];
    Mz = Mz + dt * dGasM[2];
    KinE = 0.5 * (Mx * Mx + My * My + Mz * Mz) / D;
    U[DEN] = D; U[MOMX] = Mx; U[MOMY] = My; U[MOMZ] = Mz;
    U[ENERGY] = KinE + P / (M->gamma - 1);
}

static void apply_update(const orc_dims *D, const orc_model *M, const orc_grid *G,
                         orc_state *S, orc_state *oS, orc_solver *W, double dt)
{   /* mhdgroup.F90:79-258 applyUpdate */
    long nc = D->nc, nf = D->nf;
    /* CopyState (mhdgroup.F90:204-255) */
    memcpy(oS->Gas, S->Gas, sizeof(double) * (size_t)nc * 5);
    memcpy(oS->Bxyz, S->Bxyz, sizeof(double) * (size_t)nc * 3);
    memcpy(oS->magFlux, S->magFlux, sizeof(double) * (size_t)nf * 3);
    oS->time = S->time;
    S->time = S->time + dt;
    /* Reynolds with Model%dt (mhdgroup.F90:114-125) */
#pragma omp parallel for collapse(2)
    for (int k = D->ks; k <= D->ke; k++)
        for (int j = D->js; j <= D->je; j++)
            for (int i = D->is; i <= D->ie; i++) {
                long c = CI(D, i, j, k);
                for (int v = 0; v < 5; v++) S->Gas[c + nc * v] = S->Gas[c + nc * v] + M->dt * W->dGasH[c + nc * v];
            }
    if (!M->doMHD) return;
    /* E2Flux (fields.F90:18-51) */
    const double *E = S->Efld;
#pragma omp parallel for collapse(2)
    for (int k = D->ks; k <= D->ke + 1; k++)
        for (int j = D->js; j <= D->je + 1; j++)
            for (int i = D->is; i <= D->ie + 1; i++) {
                long c = CI(D, i, j, k), f = FI(D, i, j, k);
                double dfi = E[CI(D, i, j + 1, k) + nc * KDIR] - E[c + nc * KDIR] - E[CI(D, i, j, k + 1) + nc * JDIR] + E[c + nc * JDIR];
                double dfj = E[CI(D, i, j, k + 1) + nc * IDIR] - E[c + nc * IDIR] - E[CI(D, i + 1, j, k) + nc * KDIR] + E[c + nc * KDIR];
                double dfk = E[CI(D, i + 1, j, k) + nc * JDIR] - E[c + nc * JDIR] - E[CI(D, i, j + 1, k) + nc * IDIR] + E[c + nc * IDIR];
                S->magFlux[f + nf * 0] = S->magFlux[f + nf * 0] - dt * dfi;
                S->magFlux[f + nf * 1] = S->magFlux[f + nf * 1] - dt * dfj;
                S->magFlux[f + nf * 2] = S->magFlux[f + nf * 2] - dt * dfk;
            }
    orc_bflux2fld(D, M, G, S->magFlux, S->Bxyz);
    /* Maxwell / Boris (mhdgroup.F90:171-195) */
#pragma omp parallel for collapse(2)
    for (int k = D->ks; k <= D->ke; k++)
        for (int j = D->js; j <= D->je; j++)
            for (int i = D->is; i <= D->ie; i++) {
                long c = CI(D, i, j, k);
                double U[5], oU[5], B[3], oB[3], dPg[5], dPm[3];
                for (int v = 0; v < 5; v++) { U[v] = S->Gas[c + nc * v]; oU[v] = oS->Gas[c + nc * v]; dPg[v] = W->dGasH[c + nc * v]; }
                for (int x = 0; x < 3; x++) {
                    B[x] = S->Bxyz[c + nc * x]; oB[x] = oS->Bxyz[c + nc * x]; dPm[x] = W->dGasM[c + nc * x];
                    if (M->doBackground) { B[x] = B[x] + G->B0[c + nc * x]; oB[x] = oB[x] + G->B0[c + nc * x]; }
                }
                if (M->doBoris) cell_boris(M, U, oU, B, oB, dPg, dPm, M->Ca, dt);
                else cell_maxwell(M, U, dPm, dt);
                for (int v = 0; v < 5; v++) S->Gas[c + nc * v] = U[v];
            }
}

void orc_advance_mhd(const orc_dims *D, orc_model *M, const orc_grid *G,
                     orc_state *S, orc_state *oS, orc_solver *W, double dt)
{   /* mhdgroup.F90:16-76 AdvanceMHD */
    init_weights();
    if (fabs(oS->time - S->time) < TINY) oS->time = S->time - M->dt;
    predictor(D, M, G, S, oS, &W->hf, 0.5 * dt);
    if (M->doMHD) calc_elec_field(D, M, G, S, &W->hf, W->Vf, S->Efld);
    calc_stress(D, M, G, S, &W->hf, W);
    apply_update(D, M, G, S, oS, W, dt);
}

int orc_step(const orc_dims *D, orc_model *M, const orc_grid *G,
             orc_state *S, orc_state *oS, orc_solver *W)
{   /* gamapp.F90:73-113 stepGamera + UpdateStateData */
    orc_advance_mhd(D, M, G, S, oS, W, M->dt);
    M->ts = M->ts + 1;
    M->t = M->t + M->dt;
    double dt;
    int st = orc_calc_dt(D, M, G, S, &dt);
    M->dt = dt;
    orc_enforce_bcs(D, M, G, S);
    orc_bflux2fld(D, M, G, S->magFlux, S->Bxyz);
    return st;
}

int orc_init_run(const orc_dims *D, orc_model *M, const orc_grid *G,
                 orc_state *S, orc_state *oS)
{   /* init.F90:200-244 (CalcGridInfo tail, no restart) */
    long nc = D->nc, nf = D->nf;
    if (M->doMHD) orc_bflux2fld(D, M, G, S->magFlux, S->Bxyz);
    orc_enforce_bcs(D, M, G, S);
    memcpy(oS->Gas, S->Gas, sizeof(double) * (size_t)nc * 5);
    memcpy(oS->Bxyz, S->Bxyz, sizeof(double) * (size_t)nc * 3);
    memcpy(oS->magFlux, S->magFlux, sizeof(double) * (size_t)nf * 3);
    memcpy(oS->Efld, S->Efld, sizeof(double) * (size_t)nc * 3);
    oS->time = S->time;
    double dt;
    int st = orc_calc_dt(D, M, G, S, &dt);
    M->dt = dt;
    oS->time = S->time - M->dt;
    return st;
}

double orc_divb(const orc_dims *D, const orc_model *M, const orc_grid *G,
                const orc_state *S, double *divout)
{   /* gridutils.F90:717-770 DivB (doTot) */
    long nf = D->nf, nc = D->nc;
    double tot = 0.0;
    for (int k = D->ks; k <= D->ke; k++)
        for (int j = D->js; j <= D->je; j++)
            for (int i = D->is; i <= D->ie; i++) {
                const double *mf = S->magFlux;
                double dFi = mf[FI(D, i, j, k)] - mf[FI(D, i + 1, j, k)];
                double dFj = mf[FI(D, i, j, k) + nf] - mf[FI(D, i, j + 1, k) + nf];
                double dFk = mf[FI(D, i, j, k) + 2 * nf] - mf[FI(D, i, j, k + 1) + 2 * nf];
                double dFi0 = 0.0, dFj0 = 0.0, dFk0 = 0.0;
                if (M->doBackground && G->bFlux0) {
                    dFi0 = G->bFlux0[CI(D, i, j, k)] - G->bFlux0[CI(D, i + 1, j, k)];
                    dFj0 = G->bFlux0[CI(D, i, j, k) + nc] - G->bFlux0[CI(D, i, j + 1, k) + nc];
                    dFk0 = G->bFlux0[CI(D, i, j, k) + 2 * nc] - G->bFlux0[CI(D, i, j, k + 1) + 2 * nc];
                }
                double v = fabs(dFi + dFj + dFk + dFi0 + dFj0 + dFk0);
                if (divout) divout[CI(D, i, j, k)] = v;
                tot = tot + v;
            }
    return tot;
}

/* ================================================== grid pre-compute */
static void xyz_at(const orc_dims *D, const orc_grid *G, int i, int j, int k, double *p)
{
    long f = FI(D, i, j, k);
    p[0] = G->xyz[f]; p[1] = G->xyz[f + D->nf]; p[2] = G->xyz[f + 2 * D->nf];
}

static void face_coords(const orc_dims *D, const orc_grid *G, int i, int j, int k, int d,
                        double *f0, double *f1, double *f2, double *f3)
{   /* metric.F90:75-104 faceCoords */
    xyz_at(D, G, i, j, k, f0);
    if (d == IDIR) { xyz_at(D, G, i, j + 1, k, f1); xyz_at(D, G, i, j, k + 1, f2); xyz_at(D, G, i, j + 1, k + 1, f3); }
    else if (d == JDIR) { xyz_at(D, G, i, j, k + 1, f1); xyz_at(D, G, i + 1, j, k, f2); xyz_at(D, G, i + 1, j, k + 1, f3); }
    else { xyz_at(D, G, i + 1, j, k, f1); xyz_at(D, G, i, j + 1, k, f2); xyz_at(D, G, i + 1, j + 1, k, f3); }
}

static void face_deltas(const double *f0, const double *f1, const double *f2, const double *f3,
                        double *dx, double *dy, double *dz)
{   /* quadrature.F90:371-389 faceDeltas */
    dx[0] = f1[0] - f0[0]; dx[1] = f2[0] - f0[0]; dx[2] = f3[0] + f0[0] - f2[0] - f1[0];
    dy[0] = f1[1] - f0[1]; dy[1] = f2[1] - f0[1]; dy[2] = f3[1] + f0[1] - f2[1] - f1[1];
    dz[0] = f1[2] - f0[2]; dz[1] = f2[2] - f0[2]; dz[2] = f3[2] + f0[2] - f2[2] - f1[2];
}

/* Face quadrature, covering GaussianFaceIntegral (mode 0: area and
 * centroid; mode 1: integral of B0), GaussianFaceFlux (flux) and
 * GaussianFaceStress (Mbb).  quadrature.F90:98-300 */
static void face_quad(const double *f0, const double *f1, const double *f2, const double *f3,
                      int mode, double M0, double *area, double *fInt, double *flx, double *Mbb)
{
    double dx[3], dy[3], dz[3];
    face_deltas(f0, f1, f2, f3, dx, dy, dz);
    double A = 0.0, F[3] = {0, 0, 0}, fl = 0.0, S[3] = {0, 0, 0};
    for (int a = 0; a < 12; a++)
        for (int b = 0; b < 12; b++) {
            double eta = 0.5 * (1 + gA[a]), psi = 0.5 * (1 + gA[b]);
            double w = gW[a] * gW[b];
            double xp = f0[0] + dx[0] * eta + dx[1] * psi + dx[2] * eta * psi;
            double yp = f0[1] + dy[0] * eta + dy[1] * psi + dy[2] * eta * psi;
            double zp = f0[2] + dz[0] * eta + dz[1] * psi + dz[2] * eta * psi;
            double etaV[3] = {dx[0] + dx[2] * psi, dy[0] + dy[2] * psi, dz[0] + dz[2] * psi};
            double psiV[3] = {dx[1] + dx[2] * eta, dy[1] + dy[2] * eta, dz[1] + dz[2] * eta};
            double cr[3];
            cross3(etaV, psiV, cr);
            double ar = 0.25 * norm3(cr) * w;
            if (mode == 0) {
                A = A + ar * 1.0;
                F[0] = F[0] + ar * xp; F[1] = F[1] + ar * yp; F[2] = F[2] + ar * zp;
            } else {
                double B[3];
                orc_dipole(M0, xp, yp, zp, B);
                F[0] = F[0] + ar * B[0]; F[1] = F[1] + ar * B[1]; F[2] = F[2] + ar * B[2];
                double N[3] = {0.25 * w * cr[0], 0.25 * w * cr[1], 0.25 * w * cr[2]};
                fl = fl + dot3(N, B);
                double bb = dot3(B, B), bdS = dot3(B, N);
                for (int x = 0; x < 3; x++) S[x] = S[x] + 0.5 * bb * N[x] - B[x] * bdS;
            }
        }
    if (area) *area = A;
    if (fInt) { fInt[0] = F[0]; fInt[1] = F[1]; fInt[2] = F[2]; }
    if (flx) *flx = fl;
    if (Mbb) { Mbb[0] = S[0]; Mbb[1] = S[1]; Mbb[2] = S[2]; }
}

/* GaussianVolumeIntegral (quadrature.F90:42-96) of (1,x,y,z) or B0 */
static void vol_quad(const double xyzC[8][3], int mode, double M0, double *wInt)
{
    static const int iS[8] = {-1, +1, +1, -1, -1, +1, +1, -1};
    static const int jS[8] = {-1, -1, +1, +1, -1, -1, +1, +1};
    static const int kS[8] = {-1, -1, -1, -1, +1, +1, +1, +1};
    for (int v = 0; v < 5; v++) wInt[v] = 0.0;
    for (int a = 0; a < 12; a++)
        for (int b = 0; b < 12; b++)
            for (int c = 0; c < 12; c++) {
                double wijk = gW[a] * gW[b] * gW[c];
                double iScl[8], jScl[8], kScl[8];
                for (int q = 0; q < 8; q++) {
                    iScl[q] = (iS[q] < 0) ? 1 - gA[a] : 1 + gA[a];
                    jScl[q] = (jS[q] < 0) ? 1 - gA[b] : 1 + gA[b];
                    kScl[q] = (kS[q] < 0) ? 1 - gA[c] : 1 + gA[c];
                }
                double xp[3], J[3][3];
                for (int x = 0; x < 3; x++) {
                    double s = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
                    for (int q = 0; q < 8; q++) {
                        s = s + xyzC[q][x] * (iScl[q] * jScl[q] * kScl[q]);
                        s1 = s1 + xyzC[q][x] * (iS[q] * jScl[q] * kScl[q]);
                        s2 = s2 + xyzC[q][x] * (iScl[q] * jS[q] * kScl[q]);
                        s3 = s3 + xyzC[q][x] * (iScl[q] * jScl[q] * kS[q]);
                    }
                    xp[x] = 0.125 * s;
                    J[x][0] = 0.125 * s1; J[x][1] = 0.125 * s2; J[x][2] = 0.125 * s3;
                }
                double dJ = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
                          - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
                          + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                dJ = fabs(dJ);
                double Wp[5];
                if (mode == 0) { Wp[0] = 1; Wp[1] = xp[0]; Wp[2] = xp[1]; Wp[3] = xp[2]; Wp[4] = 0; }
                else { double B[3]; orc_dipole(M0, xp[0], xp[1], xp[2], B); Wp[0] = B[0]; Wp[1] = B[1]; Wp[2] = B[2]; Wp[3] = 0; Wp[4] = 0; }
                for (int v = 0; v < 5; v++) wInt[v] = wInt[v] + dJ * wijk * Wp[v];
            }
}

static void cell_coords(const orc_dims *D, const orc_grid *G, int i, int j, int k, double xyzC[8][3])
{   /* metric.F90:127-142 cellCoords */
    xyz_at(D, G, i, j, k, xyzC[0]);
    xyz_at(D, G, i + 1, j, k, xyzC[1]);
    xyz_at(D, G, i + 1, j + 1, k, xyzC[2]);
    xyz_at(D, G, i, j + 1, k, xyzC[3]);
    xyz_at(D, G, i, j, k + 1, xyzC[4]);
    xyz_at(D, G, i + 1, j, k + 1, xyzC[5]);
    xyz_at(D, G, i + 1, j + 1, k + 1, xyzC[6]);
    xyz_at(D, G, i, j + 1, k + 1, xyzC[7]);
}

void orc_dipole(double M0, double x, double y, double z, double *B)
{   /* msphutils.F90:534-551 Dipole */
    double r[3] = {x, y, z}, m[3] = {0.0, 0.0, M0};
    double rad = norm3(r);
    double md = dot3(m, r);
    double r5 = pow(rad, 5.0), r3 = pow(rad, 3.0);
    for (int q = 0; q < 3; q++) B[q] = 3 * md * r[q] / r5 - m[q] / r3;
}

static void eb_geom(const orc_dims *D, orc_grid *G, int dNorm, int dT1, int dT2, int c0)
{   /* init.F90:1075-1137 ebGeom */
    long nf = D->nf;
    for (int k = D->ks; k <= D->ke + 1; k++)
        for (int j = D->js; j <= D->je + 1; j++)
            for (int i = D->is; i <= D->ie + 1; i++) {
                double fArB[8], Nxyz[3][8];
                for (int n = 0; n < 8; n++) {
                    int o = n - 4;
                    long f = FI(D, i + o * EIJK[dT2][0], j + o * EIJK[dT2][1], k + o * EIJK[dT2][2]);
                    fArB[n] = G->face[f + nf * dT1];
                    for (int x = 0; x < 3; x++) Nxyz[x][n] = G->Tf[f + nf * (NORMX + x + 9 * dT1)];
                }
                double fAi = dot8(W6, fArB), Nec[3];
                for (int x = 0; x < 3; x++) {
                    double pr[8];
                    for (int n = 0; n < 8; n++) pr[n] = Nxyz[x][n] * fArB[n];
                    Nec[x] = dot8(W6, pr) / fAi;
                }
                long f = FI(D, i, j, k);
                double t1[3] = {G->Te[f + nf * (TAN1X + 9 * dNorm)], G->Te[f + nf * (TAN1Y + 9 * dNorm)], G->Te[f + nf * (TAN1Z + 9 * dNorm)]};
                double t2[3] = {G->Te[f + nf * (TAN2X + 9 * dNorm)], G->Te[f + nf * (TAN2Y + 9 * dNorm)], G->Te[f + nf * (TAN2Z + 9 * dNorm)]};
                double dxT = dot3(Nec, t1), dyT = dot3(Nec, t2);
                double dScl = sqrt(dxT * dxT + dyT * dyT);
                if (dScl <= TINY) dScl = 1.0;
                G->Teb[f + nf * (c0 + 0 + 4 * dNorm)] = dxT / dScl;
                G->Teb[f + nf * (c0 + 1 + 4 * dNorm)] = dyT / dScl;
            }
}

void orc_corners2grid(const orc_dims *D, orc_grid *G)
{   /* init.F90:793-1072 Corners2Grid (doQuadFT = F) */
    init_weights();
    long nf = D->nf, nc = D->nc;
    for (long q = 0; q < nf * 3; q++) { G->face[q] = 1.0; G->edge[q] = 1.0; }
    memset(G->xfc, 0, sizeof(double) * (size_t)nf * 9);
    memset(G->Tf, 0, sizeof(double) * (size_t)nf * 27);
    memset(G->Te, 0, sizeof(double) * (size_t)nf * 27);
    for (long q = 0; q < nf * 12; q++) G->Teb[q] = 1.0;
    for (int d = 0; d < 3; d++) {
#pragma omp parallel for collapse(2)
        for (int k = D->ksg; k <= D->keg + EIJK[d][2]; k++)
            for (int j = D->jsg; j <= D->jeg + EIJK[d][1]; j++)
                for (int i = D->isg; i <= D->ieg + EIJK[d][0]; i++) {
                    double f0[3], f1[3], f2[3], f3[3], fA, fc[3];
                    face_coords(D, G, i, j, k, d, f0, f1, f2, f3);
                    face_quad(f0, f1, f2, f3, 0, 0.0, &fA, fc, 0, 0);
                    long f = FI(D, i, j, k);
                    for (int x = 0; x < 3; x++) G->xfc[f + nf * (x + 3 * d)] = fc[x] / fA;
                    G->face[f + nf * d] = fA;
                    double a[3], b[3], c[3], e[3], dT1[3], dT2[3];
                    for (int x = 0; x < 3; x++) {
                        if (d == IDIR) {
                            a[x] = G->xyz[FI(D, i, j + 1, k + 1) + nf * x]; b[x] = G->xyz[FI(D, i, j, k + 1) + nf * x];
                            c[x] = G->xyz[FI(D, i, j + 1, k) + nf * x]; e[x] = G->xyz[FI(D, i, j, k) + nf * x];
                            dT1[x] = 0.5 * (a[x] - b[x]) + 0.5 * (c[x] - e[x]);
                            dT2[x] = 0.5 * (G->xyz[FI(D, i, j + 1, k + 1) + nf * x] - G->xyz[FI(D, i, j + 1, k) + nf * x])
                                   + 0.5 * (G->xyz[FI(D, i, j, k + 1) + nf * x] - G->xyz[FI(D, i, j, k) + nf * x]);
                        } else if (d == JDIR) {
                            dT1[x] = 0.5 * (G->xyz[FI(D, i + 1, j, k + 1) + nf * x] - G->xyz[FI(D, i + 1, j, k) + nf * x])
                                   + 0.5 * (G->xyz[FI(D, i, j, k + 1) + nf * x] - G->xyz[FI(D, i, j, k) + nf * x]);
                            dT2[x] = 0.5 * (G->xyz[FI(D, i + 1, j, k + 1) + nf * x] - G->xyz[FI(D, i, j, k + 1) + nf * x])
                                   + 0.5 * (G->xyz[FI(D, i + 1, j, k) + nf * x] - G->xyz[FI(D, i, j, k) + nf * x]);
                        } else {
                            dT1[x] = 0.5 * (G->xyz[FI(D, i + 1, j + 1, k) + nf * x] - G->xyz[FI(D, i, j + 1, k) + nf * x])
                                   + 0.5 * (G->xyz[FI(D, i + 1, j, k) + nf * x] - G->xyz[FI(D, i, j, k) + nf * x]);
                            dT2[x] = 0.5 * (G->xyz[FI(D, i + 1, j + 1, k) + nf * x] - G->xyz[FI(D, i + 1, j, k) + nf * x])
                                   + 0.5 * (G->xyz[FI(D, i, j + 1, k) + nf * x] - G->xyz[FI(D, i, j, k) + nf * x]);
                        }
                    }
                    double cr[3], eN[3], eT1[3], eT2[3], c2[3];
                    cross3(dT1, dT2, cr);
                    double ncr = norm3(cr), nd2 = norm3(dT2);
                    for (int x = 0; x < 3; x++) { eN[x] = cr[x] / ncr; eT2[x] = dT2[x] / nd2; }
                    cross3(eT2, eN, c2);
                    double nc2 = norm3(c2);
                    for (int x = 0; x < 3; x++) eT1[x] = c2[x] / nc2;
                    for (int x = 0; x < 3; x++) {
                        G->Tf[f + nf * (NORMX + x + 9 * d)] = eN[x];
                        G->Tf[f + nf * (TAN1X + x + 9 * d)] = eT1[x];
                        G->Tf[f + nf * (TAN2X + x + 9 * d)] = eT2[x];
                    }
                }
    }
    /* Volumes, centroids (init.F90:911-940) and di/dj/dk (943-957) */
#pragma omp parallel for collapse(2)
    for (int k = D->ksg; k <= D->keg; k++)
        for (int j = D->jsg; j <= D->jeg; j++)
            for (int i = D->isg; i <= D->ieg; i++) {
                double xyzC[8][3], wI[5];
                cell_coords(D, G, i, j, k, xyzC);
                vol_quad((const double(*)[3])xyzC, 0, 0.0, wI);
                long c = CI(D, i, j, k);
                G->volume[c] = wI[0];
                for (int x = 0; x < 3; x++) G->xyzcc[c + nc * x] = wI[1 + x] / wI[0];
                double fa;
                fa = fmaxd(G->face[FI(D, i + 1, j, k) + 0 * nf], G->face[FI(D, i, j, k) + 0 * nf]);
                G->di[c] = G->volume[c] / fa;
                fa = fmaxd(G->face[FI(D, i, j + 1, k) + 1 * nf], G->face[FI(D, i, j, k) + 1 * nf]);
                G->dj[c] = G->volume[c] / fa;
                fa = fmaxd(G->face[FI(D, i, j, k + 1) + 2 * nf], G->face[FI(D, i, j, k) + 2 * nf]);
                G->dk[c] = G->volume[c] / fa;
            }
    /* Edge triads (init.F90:964-1035) */
    for (int d = 0; d < 3; d++) {
        int is = D->isg, ie = D->ieg + 1, js = D->jsg, je = D->jeg + 1, ks = D->ksg, ke = D->keg + 1;
        if (d == IDIR) { ie = D->ieg; ks = D->ksg + 1; ke = D->keg; }
        else if (d == JDIR) { je = D->jeg; is = D->isg + 1; ie = D->ieg; }
        else { ke = D->keg; js = D->jsg + 1; je = D->jeg; }
#pragma omp parallel for collapse(2)
        for (int k = ks; k <= ke; k++)
            for (int j = js; j <= je; j++)
                for (int i = is; i <= ie; i++) {
                    double dE[3], eA[3], p0[3], p1[3], a0[3], a1[3], b0[3], b1[3];
                    if (d == IDIR) {
                        xyz_at(D, G, i + 1, j, k, p1); xyz_at(D, G, i, j, k, p0);
                        xyz_at(D, G, i, j, k + 1, a1); xyz_at(D, G, i, j, k - 1, a0);
                        xyz_at(D, G, i + 1, j, k + 1, b1); xyz_at(D, G, i + 1, j, k - 1, b0);
                    } else if (d == JDIR) {
                        xyz_at(D, G, i, j + 1, k, p1); xyz_at(D, G, i, j, k, p0);
                        xyz_at(D, G, i + 1, j, k, a1); xyz_at(D, G, i - 1, j, k, a0);
                        xyz_at(D, G, i + 1, j + 1, k, b1); xyz_at(D, G, i - 1, j + 1, k, b0);
                    } else {
                        xyz_at(D, G, i, j, k + 1, p1); xyz_at(D, G, i, j, k, p0);
                        xyz_at(D, G, i, j + 1, k, a1); xyz_at(D, G, i, j - 1, k, a0);
                        xyz_at(D, G, i, j + 1, k + 1, b1); xyz_at(D, G, i, j - 1, k + 1, b0);
                    }
                    for (int x = 0; x < 3; x++) {
                        dE[x] = p1[x] - p0[x];
                        eA[x] = 0.5 * (a1[x] - a0[x] + b1[x] - b0[x]);
                    }
                    double ndE = norm3(dE), eN[3], c1[3], eT1[3], c2[3], eT2[3];
                    for (int x = 0; x < 3; x++) eN[x] = dE[x] / ndE;
                    cross3(eA, eN, c1);
                    double n1 = norm3(c1);
                    for (int x = 0; x < 3; x++) eT1[x] = c1[x] / n1;
                    cross3(eN, eT1, c2);
                    double n2 = norm3(c2);
                    for (int x = 0; x < 3; x++) eT2[x] = c2[x] / n2;
                    long f = FI(D, i, j, k);
                    G->edge[f + nf * d] = norm3(dE);
                    for (int x = 0; x < 3; x++) {
                        G->Te[f + nf * (NORMX + x + 9 * d)] = eN[x];
                        G->Te[f + nf * (TAN1X + x + 9 * d)] = eT1[x];
                        G->Te[f + nf * (TAN2X + x + 9 * d)] = eT2[x];
                    }
                }
    }
    /* Teb (init.F90:1038-1063) */
    for (int dN = 0; dN < 3; dN++) {
        int T1 = (dN == IDIR) ? JDIR : (dN == JDIR) ? KDIR : IDIR;
        int T2 = (dN == IDIR) ? KDIR : (dN == JDIR) ? IDIR : JDIR;
        eb_geom(D, G, dN, T1, T2, XNQI);
        eb_geom(D, G, dN, T2, T1, XNQJ);
    }
}

void orc_add_dipole_b0(const orc_dims *D, orc_grid *G, double M0)
{   /* background.F90:20-167 AddB0 with the Dipole field */
    long nc = D->nc, nf = D->nf;
    double *fs = (double *)calloc((size_t)nc * 9, sizeof(double)); /* faceStress(i,j,k,d,x) */
    memset(G->fcB0, 0, sizeof(double) * (size_t)nc * 9);
    memset(G->edgB0, 0, sizeof(double) * (size_t)nc * 6);
    memset(G->B0, 0, sizeof(double) * (size_t)nc * 3);
    memset(G->dpB0, 0, sizeof(double) * (size_t)nc * 3);
    memset(G->bFlux0, 0, sizeof(double) * (size_t)nc * 3);
#pragma omp parallel for collapse(2)
    for (int k = D->ksg; k <= D->keg; k++)
        for (int j = D->jsg; j <= D->jeg; j++)
            for (int i = D->isg; i <= D->ieg; i++) {
                long c = CI(D, i, j, k), f = FI(D, i, j, k);
                for (int d = 0; d < 3; d++) {
                    double f0[3], f1[3], f2[3], f3[3], flx, fInt[3], Mbb[3];
                    face_coords(D, G, i, j, k, d, f0, f1, f2, f3);
                    face_quad(f0, f1, f2, f3, 1, M0, 0, fInt, &flx, Mbb);
                    G->bFlux0[c + nc * d] = flx;
                    for (int x = 0; x < 3; x++) G->fcB0[c + nc * (x + 3 * d)] = fInt[x] / G->face[f + nf * d];
                    for (int x = 0; x < 3; x++) fs[c + nc * (d + 3 * x)] = Mbb[x];
                }
                for (int d = 0; d < 3; d++) {
                    double e1[3], e2[3];
                    xyz_at(D, G, i, j, k, e1);
                    xyz_at(D, G, i + EIJK[d][0], j + EIJK[d][1], k + EIJK[d][2], e2);
                    double dR2[3], xB[3], eInt[3] = {0, 0, 0};
                    for (int x = 0; x < 3; x++) { dR2[x] = 0.5 * (e2[x] - e1[x]); xB[x] = 0.5 * (e2[x] + e1[x]); }
                    for (int q = 0; q < 12; q++) {
                        double xp[3] = {xB[0] + gA[q] * dR2[0], xB[1] + gA[q] * dR2[1], xB[2] + gA[q] * dR2[2]}, Ap[3];
                        orc_dipole(M0, xp[0], xp[1], xp[2], Ap);
                        for (int x = 0; x < 3; x++) eInt[x] = eInt[x] + 0.5 * gW[q] * Ap[x];
                    }
                    const double *Te = G->Te;
                    G->edgB0[c + nc * (0 + 2 * d)] = eInt[0] * Te[f + nf * (TAN1X + 9 * d)] + eInt[1] * Te[f + nf * (TAN1Y + 9 * d)] + eInt[2] * Te[f + nf * (TAN1Z + 9 * d)];
                    G->edgB0[c + nc * (1 + 2 * d)] = eInt[0] * Te[f + nf * (TAN2X + 9 * d)] + eInt[1] * Te[f + nf * (TAN2Y + 9 * d)] + eInt[2] * Te[f + nf * (TAN2Z + 9 * d)];
                }
                double xyzC[8][3], wI[5];
                cell_coords(D, G, i, j, k, xyzC);
                vol_quad((const double(*)[3])xyzC, 1, M0, wI);
                for (int x = 0; x < 3; x++) G->B0[c + nc * x] = wI[x] / G->volume[c];
            }
#pragma omp parallel for collapse(2)
    for (int k = D->ks; k <= D->ke; k++)
        for (int j = D->js; j <= D->je; j++)
            for (int i = D->is; i <= D->ie; i++) {
                long c = CI(D, i, j, k);
                long cI = CI(D, i + 1, j, k), cJ = CI(D, i, j + 1, k), cK = CI(D, i, j, k + 1);
                for (int x = 0; x < 3; x++)
                    G->dpB0[c + nc * x] = -(fs[cI + nc * (0 + 3 * x)] - fs[c + nc * (0 + 3 * x)]
                                          + fs[cJ + nc * (1 + 3 * x)] - fs[c + nc * (1 + 3 * x)]
                                          + fs[cK + nc * (2 + 3 * x)] - fs[c + nc * (2 + 3 * x)]) / G->volume[c];
            }
    free(fs);
}

/* ------------------------------------------------------ pins
 * Thin wrappers that expose static pieces of the step to the oracle's own
 * pin tests (tests/test_oracle_pins.py).  They add no arithmetic. */
void orc_cell_boris(const orc_model *M, double *U2, const double *U0, const double *B2,
                    const double *B0, const double *dGasH, const double *dGasM, double dt)
{
    cell_boris(M, U2, U0, B2, B0, dGasH, dGasM, M->Ca, dt);
}

void orc_wind_interp(const orc_model *M, double t, double *out8)
{   /* out8 = rho, P, V(3), B(3) */
    wind_interp(M, t, &out8[0], &out8[1], &out8[2], &out8[5]);
}

void orc_wind_front(const orc_model *M, double t, double *vFr)
{
    wind_front(M, t, vFr);
}

void orc_wind_at(const orc_model *M, const double *vFr, const double *xyz, double t, double *out8)
{
    wind_at(M, vFr, xyz, t, &out8[0], &out8[1], &out8[2], &out8[5]);
}

double orc_wind_wgt(const orc_model *M, const double *vFr, const double *xyz, double t)
{
    return wind_wgt(M, vFr, xyz, t);
}

double orc_ramp_down(double r, double rC, double lC)
{
    return ramp_down(r, rC, lC);
}
