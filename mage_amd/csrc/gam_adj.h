// gam_adj.h -- hand-written reverse-mode (vector-Jacobian) products of the
// per-point functions in gam_phys.h that the hydrodynamic step uses.
//
// Each *_adj function takes the forward inputs (it recomputes what it
// needs), the adjoints of the forward outputs, and ADDS the adjoints of
// the forward inputs into the given accumulators.  Branches (floors,
// min/max, the PDM limiter's sign and clipping cases) take the branch the
// forward evaluation takes at these inputs, so each product is the exact
// derivative of the forward code wherever it is differentiable.
#pragma once
#include "gam_phys.h"

// d|x|/dx as the forward's fabs sees it (copysign keeps the sign of -0.0)
__device__ __forceinline__ double sgn1(double x) { return copysign(1.0, x); }

// ---------------------------------------------------------- conversions
// CellC2P (c2p_cell): U -> W
__device__ __forceinline__ void c2p_cell_adj(const MParm &m, const double *U, const double *aW,
                                             double *aU)
{
    const double rho = U[DEN];
    const double ir = 1.0 / rho;
    const double v[3] = {U[MOMX] * ir, U[MOMY] * ir, U[MOMZ] * ir};
    const double ke = 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double arho = (rho > m.dFloor) ? aW[DEN] : 0.0;
    double av[3] = {aW[1], aW[2], aW[3]};
    const double pr = (m.gamma - 1) * (U[ENER] - ke);
    if (pr > m.pFloor) {
        const double apr = aW[PRS];
        aU[ENER] += (m.gamma - 1) * apr;
        const double ake = -(m.gamma - 1) * apr;
        arho += 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * ake;
        for (int x = 0; x < 3; x++) av[x] += rho * v[x] * ake;
    }
    double air = 0.0;
    for (int x = 0; x < 3; x++) {
        aU[1 + x] += ir * av[x];
        air += U[1 + x] * av[x];
    }
    arho += -ir * ir * air;
    aU[DEN] += arho;
}

// CellP2C (p2c_cell): W -> U
__device__ __forceinline__ void p2c_cell_adj(const MParm &m, const double *W, const double *aU,
                                             double *aW)
{
    const double rho = dmax(W[DEN], m.dFloor);
    const double v[3] = {W[1], W[2], W[3]};
    const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    double arho = aU[DEN] + 0.5 * v2 * aU[ENER];
    for (int x = 0; x < 3; x++) {
        arho += v[x] * aU[1 + x];
        aW[1 + x] += rho * aU[1 + x] + rho * v[x] * aU[ENER];
    }
    if (W[DEN] > m.dFloor) aW[DEN] += arho;
    if (W[PRS] > m.pFloor) aW[PRS] += m.igm1 * aU[ENER];
}

// BlockCon2Prim (c2p_block): U -> W
__device__ __forceinline__ void c2p_block_adj(const MParm &m, const double *U, const double *aW,
                                              double *aU)
{
    const double d = dmax(U[DEN], m.dFloor);
    const double id = 1.0 / d;
    const double mo[3] = {U[MOMX], U[MOMY], U[MOMZ]};
    const double m2 = mo[0] * mo[0] + mo[1] * mo[1] + mo[2] * mo[2];
    const double ke = 0.5 * m2 * id;
    double ad = aW[DEN];
    double am[3], aid = 0.0;
    for (int x = 0; x < 3; x++) {
        am[x] = id * aW[1 + x];
        aid += mo[x] * aW[1 + x];
    }
    if ((m.gamma - 1) * (U[ENER] - ke) > m.pFloor) {
        aU[ENER] += (m.gamma - 1) * aW[PRS];
        const double ake = -(m.gamma - 1) * aW[PRS];
        for (int x = 0; x < 3; x++) am[x] += mo[x] * id * ake;
        aid += 0.5 * m2 * ake;
    }
    ad += -id * id * aid;
    for (int x = 0; x < 3; x++) aU[1 + x] += am[x];
    if (U[DEN] > m.dFloor) aU[DEN] += ad;
}

// ---------------------------------------------------------- limiter
// qN = dmax(dmin(a, b), dmin(qI, dmax(a, b))): adjoint of the clipped value
__device__ __forceinline__ void clip_adj(double a, double b, double qI, double aq, double &aa,
                                         double &ab, double &aqI)
{
    const double m1 = dmin(a, b), M = dmax(a, b), m2 = dmin(qI, M);
    if (m1 > m2) {
        if (a < b) aa += aq; else ab += aq;
    } else {
        if (qI < M) aqI += aq;
        else if (a > b) aa += aq;
        else ab += aq;
    }
}

// PDM (pdm_one)
__device__ __forceinline__ void pdm_one_adj(double b, double q0, double q1, double q2, double qI,
                                            double aout, double &aq0, double &aq1, double &aq2,
                                            double &aqI)
{
    const double qN = dmax(dmin(q1, q2), dmin(qI, dmax(q1, q2)));
    const double dq0 = b * (q1 - q0), dq1 = b * (q2 - q1);
    const double s1 = copysign(1.0, dq1);
    const double s01 = fabs(copysign(1.0, dq0) + s1);
    const double A = fabs(qN - q1) - s01 * fabs(dq0);
    double aqN = aout;
    if (!(0.0 > A)) {
        const double aA = -s1 * aout;
        const double sg = sgn1(qN - q1);
        aqN += sg * aA;
        aq1 -= sg * aA;
        const double adq0 = -s01 * sgn1(dq0) * aA;
        aq1 += b * adq0;
        aq0 -= b * adq0;
    }
    clip_adj(q1, q2, qI, aqN, aq1, aq2, aqI);
}

// pdmLR (pdm_pair)
__device__ __forceinline__ void pdm_pair_adj(double b, double v0, double v1, double v2, double v3,
                                             double qi, double aql, double aqr, double &av0,
                                             double &av1, double &av2, double &av3, double &aqi)
{
    const double vN = dmax(dmin(v1, v2), dmin(qi, dmax(v1, v2)));
    const double dv0 = b * (v1 - v0), dv1 = b * (v2 - v1), dv2 = b * (v3 - v2);
    const double s0 = copysign(1.0, dv0), s1 = copysign(1.0, dv1), s2 = copysign(1.0, dv2);
    const double a0 = fabs(s0 + s1), a1 = fabs(s1 + s2);
    double avN = aql + aqr;
    const double Al = fabs(vN - v1) - a0 * fabs(dv0);
    if (!(0.0 > Al)) {
        const double aA = -s1 * aql;
        const double sg = sgn1(vN - v1);
        avN += sg * aA;
        av1 -= sg * aA;
        const double adv0 = -a0 * sgn1(dv0) * aA;
        av1 += b * adv0;
        av0 -= b * adv0;
    }
    const double Ar = fabs(v2 - vN) - a1 * fabs(dv2);
    if (!(0.0 > Ar)) {
        const double aA = s1 * aqr;
        const double sg = sgn1(v2 - vN);
        av2 += sg * aA;
        avN -= sg * aA;
        const double adv2 = -a1 * sgn1(dv2) * aA;
        av3 += b * adv2;
        av2 -= b * adv2;
    }
    clip_adj(v1, v2, qi, avN, av1, av2, aqi);
}

// recon_lr_w: qv[8] volume-weighted stencil values, nq[4] limiter values
template <int RM>
__device__ __forceinline__ void recon_lr_w_adj(double pdmb, double rVi, const double *nq,
                                               const double *qv, double aql, double aqr,
                                               double *anq, double *aqv)
{
    if (RM == 1) {
        double l = qv[0] * UP7(0) + qv[1] * UP7(1) + qv[2] * UP7(2) + qv[3] * UP7(3)
                 + qv[4] * UP7(4) + qv[5] * UP7(5) + qv[6] * UP7(6);
        double r = qv[7] * UP7(0) + qv[6] * UP7(1) + qv[5] * UP7(2) + qv[4] * UP7(3)
                 + qv[3] * UP7(4) + qv[2] * UP7(5) + qv[1] * UP7(6);
        double al = 0.0, ar = 0.0;
        pdm_one_adj(pdmb, nq[0], nq[1], nq[2], l * rVi, aql, anq[0], anq[1], anq[2], al);
        pdm_one_adj(pdmb, nq[3], nq[2], nq[1], r * rVi, aqr, anq[3], anq[2], anq[1], ar);
        for (int n = 0; n < 7; n++) {
            aqv[n] += UP7(n) * rVi * al;
            aqv[7 - n] += UP7(n) * rVi * ar;
        }
    } else {
        double s = 0.0;
        for (int n = 0; n < 8; n++) s += W8(n) * qv[n];
        const double qi = s * rVi;
        double ai = 0.0;
        pdm_pair_adj(pdmb, nq[0], nq[1], nq[2], nq[3], qi, aql, aqr, anq[0], anq[1], anq[2],
                     anq[3], ai);
        for (int n = 0; n < 8; n++) aqv[n] += W8(n) * rVi * ai;
    }
}

// ---------------------------------------------------------- gas flux
// One side of GasKinFlux (gas_kin) without the HOGS difference dW:
// aF = adjoint of the flux contribution gF (5), adds into aW (5).
__device__ __forceinline__ void gas_kin_adj(double igm1, double s, const double *W, const double *T,
                                            const double *aF, double *aW)
{
    const double d = W[DEN], p = W[PRS], vx = W[1], vy = W[2], vz = W[3];
    const double e = 0.5 * d * (vx * vx + vy * vy + vz * vz) + p * igm1;
    const double lam = d / (2 * p);
    const double rl = rsqrt(lam);
    const double sl = lam * rl;   // sqrt(lam)
    const double vn = T[0] * vx + T[1] * vy + T[2] * vz;
    const double vt1 = T[3] * vx + T[4] * vy + T[5] * vz;
    const double vt2 = T[6] * vx + T[7] * vy + T[8] * vz;
    const double z = s * sl * vn;
    double ez;
    const double vn0 = 0.5 * gam_erfc_exp(z, ez);
    const double vn1 = vn * vn0 - 0.5 * s * ez * (rl * GAM_RSQRT_PI);
    // flux = (d vn1, T.(fn, ft1, ft2), (e + p/2) vn1 + p vn0 vn / 2)
    const double afn = T[0] * aF[MOMX] + T[1] * aF[MOMY] + T[2] * aF[MOMZ];
    const double aft1 = T[3] * aF[MOMX] + T[4] * aF[MOMY] + T[5] * aF[MOMZ];
    const double aft2 = T[6] * aF[MOMX] + T[7] * aF[MOMY] + T[8] * aF[MOMZ];
    const double aE = aF[ENER];
    double avn1 = d * aF[DEN] + (e + 0.5 * p) * aE + d * vn * afn + d * vt1 * aft1 + d * vt2 * aft2;
    double avn0 = 0.5 * p * vn * aE + p * afn;
    double avn = 0.5 * p * vn0 * aE + d * vn1 * afn;
    const double ae = vn1 * aE;
    double ap = 0.5 * vn1 * aE + 0.5 * vn0 * vn * aE + vn0 * afn;
    double ad = vn1 * aF[DEN] + vn * vn1 * afn + vt1 * vn1 * aft1 + vt2 * vn1 * aft2;
    const double avt1 = d * vn1 * aft1, avt2 = d * vn1 * aft2;
    // vn1 = vn vn0 - s ez rl / (2 sqrt(pi))
    avn += vn0 * avn1;
    avn0 += vn * avn1;
    const double aez = -0.5 * s * rl * GAM_RSQRT_PI * avn1;
    const double arl = -0.5 * s * ez * GAM_RSQRT_PI * avn1;
    // vn0 = erfc(z)/2, ez = exp(-z^2)
    const double az = -GAM_RSQRT_PI * ez * avn0 - 2.0 * z * ez * aez;
    // z = s sqrt(lam) vn
    avn += s * sl * az;
    const double asl = s * vn * az;
    const double alam = asl * 0.5 * rl - 0.5 * rl / lam * arl;
    // lam = d / (2 p)
    ad += alam / (2 * p);
    ap += -lam / p * alam;
    // e = d |v|^2 / 2 + p igm1
    ad += 0.5 * (vx * vx + vy * vy + vz * vz) * ae;
    ap += igm1 * ae;
    aW[DEN] += ad;
    aW[PRS] += ap;
    aW[1] += d * vx * ae + T[0] * avn + T[3] * avt1 + T[6] * avt2;
    aW[2] += d * vy * ae + T[1] * avn + T[4] * avt1 + T[7] * avt2;
    aW[3] += d * vz * ae + T[2] * avn + T[5] * avt1 + T[8] * avt2;
}

// ---------------------------------------------------------- predictor
// PredictCell (gas part of k_predict) at one cell: Q = State gas, oQ =
// oState gas, w = AB2 weight, vol = cell volume.  Adjoints of the
// predictor's outputs: aHG (half-state conserved, H_GAS), aWH (V * half
// primitives, H_WH), aWN (limiter primitives, H_WN).  Adds into aQ, aoQ.
__device__ __forceinline__ void predict_cell_adj(const MParm &m, const double *Q, const double *oQ,
                                                 double w, double vol, const double *aHG,
                                                 const double *aWH, const double *aWN, double *aQ,
                                                 double *aoQ)
{
    double W[5], oW[5], pW[5], pQ[5];
    c2p_cell(m, oQ, oW);
    c2p_cell(m, Q, W);
    for (int v = 0; v < 5; v++) pW[v] = ab2(W[v], oW[v], w);
    const bool dOk = pW[DEN] > m.dFloor, pOk = pW[PRS] > m.pFloor;
    pW[DEN] = dmax(pW[DEN], m.dFloor);
    pW[PRS] = dmax(pW[PRS], m.pFloor);
    p2c_cell(m, pW, pQ);
    double apQ[5], aWb[5], apW[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int v = 0; v < 5; v++) { apQ[v] = aHG[v]; aWb[v] = vol * aWH[v]; }
    if (!m.doLFM)
        for (int v = 0; v < 5; v++) aWb[v] += aWN[v];
    c2p_block_adj(m, pQ, aWb, apQ);
    p2c_cell_adj(m, pW, apQ, apW);
    if (!dOk) apW[DEN] = 0.0;
    if (!pOk) apW[PRS] = 0.0;
    double aW[5], aoW[5];
    for (int v = 0; v < 5; v++) {   // ab2: q + w (q - oq)
        aW[v] = (1.0 + w) * apW[v];
        aoW[v] = -w * apW[v];
    }
    c2p_cell_adj(m, Q, aW, aQ);
    c2p_cell_adj(m, oQ, aoW, aoQ);
    if (m.doLFM) c2p_block_adj(m, Q, aWN, aQ);
}
