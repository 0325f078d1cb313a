// gam_phys.h -- per-point FP64 physics of the GAMERA step, written for
// device code.  Each function states the reference routine it reproduces;
// operation order follows the reference so that results agree with it to
// rounding.
#pragma once
#include "gam_common.h"
#include "gam_erfc.h"

#define GAM_PI 3.141592653589793115997963468544185161590576171875
#define GAM_RSQRT_PI 0.56418958354775628694807945156077258584405062932900  // 1/sqrt(pi)

enum { DEN = 0, MOMX = 1, MOMY = 2, MOMZ = 3, ENER = 4 };
enum { PRS = 4 };

// ---------------------------------------------------------- conversions
// Divisions by a common divisor are done as one reciprocal and products
// throughout this file (last-bit differences from the reference's
// per-component divisions, well inside the FP64 parity tolerance).

// CellC2P (gamutils.F90:301-321): velocity from the unfloored density.
__device__ __forceinline__ void c2p_cell(const MParm &m, const double *U, double *W)
{
    double rho = U[DEN];
    double ir = 1.0 / rho;
    double vx = U[MOMX] * ir, vy = U[MOMY] * ir, vz = U[MOMZ] * ir;
    double ke = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
    W[DEN] = dmax(rho, m.dFloor);
    W[1] = vx; W[2] = vy; W[3] = vz;
    W[PRS] = dmax((m.gamma - 1) * (U[ENER] - ke), m.pFloor);
}

// CellP2C (gamutils.F90:362-383)
__device__ __forceinline__ void p2c_cell(const MParm &m, const double *W, double *U)
{
    double rho = dmax(W[DEN], m.dFloor);
    double vx = W[1], vy = W[2], vz = W[3];
    double ie = dmax(W[PRS], m.pFloor) * m.igm1;
    double ke = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
    U[DEN] = rho; U[MOMX] = rho * vx; U[MOMY] = rho * vy; U[MOMZ] = rho * vz;
    U[ENER] = ie + ke;
}

// The MHD side of the inflow BC's ghost gas (WindBC, wind.F90:226-335):
// CellC2P of the last active cell U (cell B Bi) with the velocity's
// component along the boundary face normal nH raised to the fast speed
// Cx = sqrt(Cs^2 + Va^2) where it is below it (Con2Fast,
// gamutils.F90:338-360; Va Boris-limited with doBoris).
__device__ __forceinline__ void wind_inner(const MParm &m, const double *U, const double *Bi, const double *nH,
                                           double *gi)
{
    c2p_cell(m, U, gi);
    const double rho = dmax(U[DEN], m.dFloor);
    const double vx = U[MOMX] / rho, vy = U[MOMY] / rho, vz = U[MOMZ] / rho;
    const double ke = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
    const double P = dmax((m.gamma - 1) * (U[ENER] - ke), m.pFloor);
    const double Cs = sqrt(m.gamma * P / rho);
    double Va = sqrt(Bi[0] * Bi[0] + Bi[1] * Bi[1] + Bi[2] * Bi[2]) / sqrt(U[DEN]);
    if (m.doBoris) Va = m.Ca * Va / sqrt(m.Ca * m.Ca + Va * Va);
    const double Cx = sqrt(Cs * Cs + Va * Va);
    double V[3] = {gi[1], gi[2], gi[3]};
    const double vp = V[0] * nH[0] + V[1] * nH[1] + V[2] * nH[2];
    if (vp < 1.0 * Cx) {
        for (int x = 0; x < 3; x++) V[x] = V[x] - vp * nH[x] + 1.0 * Cx * nH[x];
    }
    gi[1] = V[0]; gi[2] = V[1]; gi[3] = V[2];
}

// BlockCon2Prim (recon.F90:75-90): density floored before dividing.
__device__ __forceinline__ void c2p_block(const MParm &m, const double *U, double *W)
{
    double d = dmax(U[DEN], m.dFloor);
    double id = 1.0 / d;
    double mx = U[MOMX], my = U[MOMY], mz = U[MOMZ];
    double ke = 0.5 * (mx * mx + my * my + mz * mz) * id;
    W[DEN] = d; W[1] = mx * id; W[2] = my * id; W[3] = mz * id;
    W[PRS] = dmax((m.gamma - 1) * (U[ENER] - ke), m.pFloor);
}

// AB2 extrapolation (ExtrapAB2, mhdgroup.F90:709-717): q + pdt*(q-oq)/odt
// with w = pdt/odt formed once per thread.
__device__ __forceinline__ double ab2(double q, double oq, double w)
{
    return q + w * (q - oq);
}

// ---------------------------------------------------------- limiter
// PDM (recon.F90:275-301)
__device__ __forceinline__ double pdm_one(double b, double q0, double q1, double q2, double qI)
{
    double qN = dmax(dmin(q1, q2), dmin(qI, dmax(q1, q2)));
    double dq0 = b * (q1 - q0), dq1 = b * (q2 - q1);
    double s1 = copysign(1.0, dq1);
    double s01 = fabs(copysign(1.0, dq0) + s1);
    return qN - s1 * dmax(0.0, fabs(qN - q1) - s01 * fabs(dq0));
}

// pdmLR (recon.F90:231-272)
__device__ __forceinline__ void pdm_pair(double b, double v0, double v1, double v2, double v3,
                                         double qi, double &ql, double &qr)
{
    double vN = dmax(dmin(v1, v2), dmin(qi, dmax(v1, v2)));
    double dv0 = b * (v1 - v0), dv1 = b * (v2 - v1), dv2 = b * (v3 - v2);
    double s0 = copysign(1.0, dv0), s1 = copysign(1.0, dv1), s2 = copysign(1.0, dv2);
    double a0 = fabs(s0 + s1), a1 = fabs(s1 + s2);
    ql = vN - s1 * dmax(0.0, fabs(vN - v1) - a0 * fabs(dv0));
    qr = vN + s1 * dmax(0.0, fabs(v2 - vN) - a1 * fabs(dv2));
}

// Volume-weighted reconstruction + limiting for one variable at one face:
// Cen8LRs (recon.F90:125-158) or Up7LRs (162-198).  vol[8] the weights
// (cells i-4..i+3), Vi their central-8 interpolant, q[8] the values,
// nq[4] the limiter values (cells i-2..i+1).  rVi = 1/Vi.
template <int RM>
__device__ __forceinline__ void recon_lr(double pdmb, const double *vol, double rVi,
                                         const double *nq, const double *q,
                                         double &ql, double &qr)
{
    double qv[8];
#pragma unroll
    for (int n = 0; n < 8; n++) qv[n] = vol[n] * q[n];
    if (RM == 1) {
        double l = qv[0] * UP7(0) + qv[1] * UP7(1) + qv[2] * UP7(2) + qv[3] * UP7(3)
                 + qv[4] * UP7(4) + qv[5] * UP7(5) + qv[6] * UP7(6);
        double r = qv[7] * UP7(0) + qv[6] * UP7(1) + qv[5] * UP7(2) + qv[4] * UP7(3)
                 + qv[3] * UP7(4) + qv[2] * UP7(5) + qv[1] * UP7(6);
        ql = pdm_one(pdmb, nq[0], nq[1], nq[2], l * rVi);
        qr = pdm_one(pdmb, nq[3], nq[2], nq[1], r * rVi);
    } else {
        double qi = wsum8([&](int n) { return qv[n]; }) * rVi;
        pdm_pair(pdmb, nq[0], nq[1], nq[2], nq[3], qi, ql, qr);
    }
}

// recon_lr on values already multiplied by their weights: qv[n] = V_n q_n.
template <int RM>
__device__ __forceinline__ void recon_lr_w(double pdmb, double rVi, const double *nq,
                                           const double *qv, double &ql, double &qr)
{
    if (RM == 1) {
        double l = qv[0] * UP7(0) + qv[1] * UP7(1) + qv[2] * UP7(2) + qv[3] * UP7(3)
                 + qv[4] * UP7(4) + qv[5] * UP7(5) + qv[6] * UP7(6);
        double r = qv[7] * UP7(0) + qv[6] * UP7(1) + qv[5] * UP7(2) + qv[4] * UP7(3)
                 + qv[3] * UP7(4) + qv[2] * UP7(5) + qv[1] * UP7(6);
        ql = pdm_one(pdmb, nq[0], nq[1], nq[2], l * rVi);
        qr = pdm_one(pdmb, nq[3], nq[2], nq[1], r * rVi);
    } else {
        double qi = wsum8([&](int n) { return qv[n]; }) * rVi;
        pdm_pair(pdmb, nq[0], nq[1], nq[2], nq[3], qi, ql, qr);
    }
}

// Mom2Face (fields.F90:235-285) of one face from the volume-weighted
// velocity stencil qv[n] = V_n v_n and the stencil's volume interpolant
// Vi (rVi = 1/Vi): <V v>_8 / max(<V>_8, TINY).
__device__ __forceinline__ double mom2face(const double *qv, double Vi, double rVi)
{
    const double r = Vi > GAM_TINY ? rVi : 1.0 / GAM_TINY;
    return wsum8([&](int n) { return qv[n]; }) * r;
}

// ---------------------------------------------------------- fluxes
// One side (s = -1 left, +1 right) of the Lyon gas-kinetic flux,
// GasKinFlux (stress.F90:432-511).  T = face triad (N, T1, T2).
// sqrt(lam) and 1/sqrt(pi*lam) come from one reciprocal square root.
__device__ __forceinline__ void gas_kin(double igm1, double s, const double *W, const double *T,
                                        double *gF, double *dW)
{
    double d = W[DEN], p = W[PRS], vx = W[1], vy = W[2], vz = W[3];
    double e = 0.5 * d * (vx * vx + vy * vy + vz * vz) + p * igm1;
    double lam = d / (2 * p);
    double rl = rsqrt(lam);
    double vn = T[0] * vx + T[1] * vy + T[2] * vz;
    double vt1 = T[3] * vx + T[4] * vy + T[5] * vz;
    double vt2 = T[6] * vx + T[7] * vy + T[8] * vz;
    // erfc(s sqrt(lam) vn) and exp(-lam vn^2) from one evaluation
    double ez;
    double vn0 = 0.5 * gam_erfc_exp(s * (lam * rl) * vn, ez);
    double vn1 = vn * vn0 - 0.5 * s * ez * (rl * GAM_RSQRT_PI);
    gF[DEN] = gF[DEN] + d * vn1;
    gF[ENER] = gF[ENER] + (e + 0.5 * p) * vn1 + 0.5 * p * vn0 * vn;
    double fn = d * vn * vn1 + p * vn0;
    double ft1 = d * vt1 * vn1, ft2 = d * vt2 * vn1;
    gF[MOMX] = gF[MOMX] + (T[0] * fn + T[3] * ft1 + T[6] * ft2);
    gF[MOMY] = gF[MOMY] + (T[1] * fn + T[4] * ft1 + T[7] * ft2);
    gF[MOMZ] = gF[MOMZ] + (T[2] * fn + T[5] * ft1 + T[8] * ft2);
    dW[DEN] = dW[DEN] + s * d;
    dW[MOMX] = dW[MOMX] + s * d * vx;
    dW[MOMY] = dW[MOMY] + s * d * vy;
    dW[MOMZ] = dW[MOMZ] + s * d * vz;
    dW[ENER] = dW[ENER] + s * e;
}

// One side of the magnetic kinetic flux, MagKinFlux (stress.F90:513-599),
// with the background-field cross terms (589-596).
template <bool BG, bool BORIS>
__device__ __forceinline__ void mag_kin(const MParm &m, double s, const double *W, const double *B,
                                        double Bn, const double *T, const double *B0, double B0n,
                                        double *mF, double &bbD, double &VaD)
{
    double d = W[DEN], p = W[PRS];
    double id = 1.0 / d;
    double bx = B[0], by = B[1], bz = B[2];
    double dPb = 0.5 * (bx * bx + by * by + bz * bz);
    double va2;
    if (BG) {
        double tx = bx + B0[0], ty = by + B0[1], tz = bz + B0[2];
        double tt = tx * tx + ty * ty + tz * tz;
        va2 = tt * id;
        bbD = bbD + 0.5 * tt;
    } else {
        va2 = 2 * dPb * id;
        bbD = bbD + 0.5 * (bx * bx + by * by + bz * bz);
    }
    double vac2 = BORIS ? va2 / (1 + va2 * m.iCa2) : va2;
    double nx = T[0], ny = T[1], nz = T[2];
    double vn = nx * W[1] + ny * W[2] + nz * W[3];
    double lam = 1 / (2 * p * id + vac2);
    double ez;
    double vn0 = 0.5 * gam_erfc_exp(s * sqrt(lam) * vn, ez);
    VaD = VaD + 0.5 * sqrt(vac2);
    mF[0] = mF[0] + vn0 * (dPb * nx - bx * Bn);
    mF[1] = mF[1] + vn0 * (dPb * ny - by * Bn);
    mF[2] = mF[2] + vn0 * (dPb * nz - bz * Bn);
    if (BG) {
        double bdb0 = B0[0] * bx + B0[1] * by + B0[2] * bz;
        mF[0] = mF[0] + vn0 * (-(B0[0] * Bn) - bx * B0n + bdb0 * nx);
        mF[1] = mF[1] + vn0 * (-(B0[1] * Bn) - by * B0n + bdb0 * ny);
        mF[2] = mF[2] + vn0 * (-(B0[2] * Bn) - bz * B0n + bdb0 * nz);
    }
}

// ---------------------------------------------------------- cell B
// gamCellBxyz (gridutils.F90:648-680) given the six face fluxes.
__device__ __forceinline__ void cell_b(const Geo &g, const double *met, long c,
                                       double fI, double fJ, double fK,
                                       double fIp, double fJp, double fKp, double *B)
{
    const long N3 = g.N3;
    const unsigned oc = (unsigned)(c * 8), oI = oc + 8, oJ = oc + 8 * g.P,
                   oK = oc + (unsigned)(8 * g.PP);
    double idV = 1.0 / ldp(met + M_VOL * N3, oc);
    double div = fIp - fI + fJp - fJ + fKp - fK;
#pragma unroll
    for (int x = 0; x < 3; x++) {
        double xI = ldp(met + (M_XFC + x + 0) * N3, oc);
        double xJ = ldp(met + (M_XFC + x + 3) * N3, oc);
        double xK = ldp(met + (M_XFC + x + 6) * N3, oc);
        double xIp = ldp(met + (M_XFC + x + 0) * N3, oI);
        double xJp = ldp(met + (M_XFC + x + 3) * N3, oJ);
        double xKp = ldp(met + (M_XFC + x + 6) * N3, oK);
        double b = (fIp * xIp + fJp * xJp + fKp * xKp - fI * xI - fJ * xJ - fK * xK) * idV;
        B[x] = b - div * ldp(met + (M_XYZCC + x) * N3, oc) * idV;
    }
}

// ---------------------------------------------------------- timestep
// CellDT (gridutils.F90:49-106) on conserved U and total B.
template <bool BORIS>
__device__ __forceinline__ double cell_dt(const MParm &m, const double *U, double bx, double by,
                                          double bz, double dl)
{
    double rho = U[DEN];
    double ir = 1.0 / rho;
    double vx = U[MOMX] * ir, vy = U[MOMY] * ir, vz = U[MOMZ] * ir;
    double mv = sqrt(vx * vx + vy * vy + vz * vz);
    double ke = 0.5 * rho * (mv * mv);
    double p = (m.gamma - 1) * (U[ENER] - ke);
    double cs = sqrt(m.gamma * p * ir);
    double va = 0.0;
    if (m.doMHD) {
        va = sqrt(bx * bx + by * by + bz * bz) / sqrt(rho);
        if (BORIS) va = m.Ca * va / sqrt(m.Ca * m.Ca + va * va);
    }
    double vcfl = mv + sqrt(cs * cs + va * va) + 0.0;
    return m.CFL * dl / vcfl;
}
