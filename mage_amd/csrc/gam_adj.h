// gam_adj.h -- hand-written reverse-mode (vector-Jacobian) products of the
// per-point functions in gam_phys.h that the step uses.
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

// wind_inner: (U, Bi) -> gi, the face normal nH held fixed
__device__ __forceinline__ void wind_inner_adj(const MParm &m, const double *U, const double *Bi,
                                               const double *nH, const double *agi, double *aU, double *aB)
{
    double W[5];
    c2p_cell(m, U, W);
    const double rho = dmax(U[DEN], m.dFloor);
    const double vv[3] = {U[MOMX] / rho, U[MOMY] / rho, U[MOMZ] / rho};
    const double v2 = vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2];
    const double ke = 0.5 * rho * v2;
    const double pr = (m.gamma - 1) * (U[ENER] - ke);
    const double P = dmax(pr, m.pFloor);
    const double Cs = sqrt(m.gamma * P / rho);
    const double bm = sqrt(Bi[0] * Bi[0] + Bi[1] * Bi[1] + Bi[2] * Bi[2]);
    const double Va0 = bm / sqrt(U[DEN]);
    const double Va = m.doBoris ? m.Ca * Va0 / sqrt(m.Ca * m.Ca + Va0 * Va0) : Va0;
    const double Cx = sqrt(Cs * Cs + Va * Va);
    const double vp = W[1] * nH[0] + W[2] * nH[1] + W[3] * nH[2];
    double aW[5] = {agi[0], agi[1], agi[2], agi[3], agi[4]};
    if (vp < 1.0 * Cx) {
        // V = V0 + (Cx - V0 . nH) nH
        const double aCx = agi[1] * nH[0] + agi[2] * nH[1] + agi[3] * nH[2];
        for (int x = 0; x < 3; x++) aW[1 + x] -= aCx * nH[x];
        const double aCs = aCx * Cs / Cx, aVa = aCx * Va / Cx;
        // Cs = sqrt(gamma P / rho) with the Con2Fast rho, velocities, P
        const double aP = aCs * 0.5 * Cs / P;
        double arho = -aCs * 0.5 * Cs / rho;
        if (pr > m.pFloor) {
            aU[ENER] += (m.gamma - 1) * aP;
            const double ake = -(m.gamma - 1) * aP;
            arho += 0.5 * v2 * ake;
            double air = 0.0;
            for (int x = 0; x < 3; x++) {
                const double av = rho * vv[x] * ake;
                aU[1 + x] += av / rho;
                air += U[1 + x] * av;
            }
            arho -= air / (rho * rho);
        }
        if (U[DEN] > m.dFloor) aU[DEN] += arho;
        // Va = |B| / sqrt(rho), Boris-limited: Ca Va0 / sqrt(Ca^2 + Va0^2)
        const double q2 = m.Ca * m.Ca + Va0 * Va0;
        const double aVa0 = m.doBoris ? aVa * m.Ca * m.Ca * m.Ca / (q2 * sqrt(q2)) : aVa;
        if (bm > 0.0) {
            const double sr = sqrt(U[DEN]);
            for (int x = 0; x < 3; x++) aB[x] += aVa0 * Bi[x] / (bm * sr);
            aU[DEN] += -0.5 * aVa0 * Va0 / U[DEN];
        }
    }
    c2p_cell_adj(m, U, aW, aU);
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
// qN = dmax(dmin(a, b), dmin(qI, dmax(a, b))): adjoint of the clipped value.
// Each output takes aq or 0 by a value select: with branches the compiler
// merges the two adds into one add through a selected address, and the
// caller's arrays (the limiter adjoints) then live in scratch memory.
__device__ __forceinline__ void clip_adj(double a, double b, double qI, double aq, double &aa,
                                         double &ab, double &aqI)
{
    const double m1 = dmin(a, b), M = dmax(a, b), m2 = dmin(qI, M);
    const bool lo = m1 > m2;
    const bool toQ = !lo && qI < M;
    const bool toA = lo ? (a < b) : (!toQ && a > b);
    const bool toB = !toQ && !toA;
    aa += toA ? aq : 0.0;
    ab += toB ? aq : 0.0;
    aqI += toQ ? aq : 0.0;
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

// recon_lr_w's transpose up to its linear stencil: sl, sr = the adjoints of
// the left and right high-order values times rVi, so that the stencil
// adjoint is aqv[n] += UP7(n) sl (n <= 6) + UP7(7 - n) sr (n >= 1) for
// RM == 1, aqv[n] += W8(n) sl (sr = 0) for RM == 0; adds into anq.
template <int RM>
__device__ __forceinline__ void recon_lr_w_adj_s(double pdmb, double rVi, const double *nq,
                                                 const double *qv, double aql, double aqr,
                                                 double *anq, double &sl, double &sr)
{
    if (RM == 1) {
        double l = qv[0] * UP7(0) + qv[1] * UP7(1) + qv[2] * UP7(2) + qv[3] * UP7(3)
                 + qv[4] * UP7(4) + qv[5] * UP7(5) + qv[6] * UP7(6);
        double r = qv[7] * UP7(0) + qv[6] * UP7(1) + qv[5] * UP7(2) + qv[4] * UP7(3)
                 + qv[3] * UP7(4) + qv[2] * UP7(5) + qv[1] * UP7(6);
        double al = 0.0, ar = 0.0;
        pdm_one_adj(pdmb, nq[0], nq[1], nq[2], l * rVi, aql, anq[0], anq[1], anq[2], al);
        pdm_one_adj(pdmb, nq[3], nq[2], nq[1], r * rVi, aqr, anq[3], anq[2], anq[1], ar);
        sl = rVi * al;
        sr = rVi * ar;
    } else {
        double t = 0.0;
        for (int n = 0; n < 8; n++) t += W8(n) * qv[n];
        const double qi = t * rVi;
        double ai = 0.0;
        pdm_pair_adj(pdmb, nq[0], nq[1], nq[2], nq[3], qi, aql, aqr, anq[0], anq[1], anq[2],
                     anq[3], ai);
        sl = rVi * ai;
        sr = 0.0;
    }
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
// One side of GasKinFlux (gas_kin): aF = adjoint of the flux contribution
// gF (5), adW (may be null) that of the HOGS difference dW (5); adds into
// aW (5).
__device__ __forceinline__ void gas_kin_adj(double igm1, double s, const double *W, const double *T,
                                            const double *aF, double *aW, const double *adW = nullptr)
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
    double ae = vn1 * aE;
    double ap = 0.5 * vn1 * aE + 0.5 * vn0 * vn * aE + vn0 * afn;
    double ad = vn1 * aF[DEN] + vn * vn1 * afn + vt1 * vn1 * aft1 + vt2 * vn1 * aft2;
    const double avt1 = d * vn1 * aft1, avt2 = d * vn1 * aft2;
    double ae2 = 0.0, avx = 0.0, avy = 0.0, avz = 0.0;
    if (adW) {   // dW += s (d, d v, e)
        ad += s * (adW[DEN] + vx * adW[MOMX] + vy * adW[MOMY] + vz * adW[MOMZ]);
        avx = s * d * adW[MOMX];
        avy = s * d * adW[MOMY];
        avz = s * d * adW[MOMZ];
        ae2 = s * adW[ENER];
    }
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
    ae += ae2;
    ad += 0.5 * (vx * vx + vy * vy + vz * vz) * ae;
    ap += igm1 * ae;
    aW[DEN] += ad;
    aW[PRS] += ap;
    aW[1] += d * vx * ae + T[0] * avn + T[3] * avt1 + T[6] * avt2 + avx;
    aW[2] += d * vy * ae + T[1] * avn + T[4] * avt1 + T[7] * avt2 + avy;
    aW[3] += d * vz * ae + T[2] * avn + T[5] * avt1 + T[8] * avt2 + avz;
}

// ---------------------------------------------------------- magnetic flux
// One side of MagKinFlux (mag_kin<BG, BORIS>): aF = adjoint of the flux
// contribution mF (3), aVaD and abbD those of the side's VaD and bbD
// terms.  Adds into aW (rho, P, v), aB (3) and aBn; B0, B0n are constants.
template <bool BG, bool BORIS>
__device__ __forceinline__ void mag_kin_adj(const MParm &m, double s, const double *W, const double *B,
                                            double Bn, const double *T, const double *B0, double B0n,
                                            const double *aF, double aVaD, double abbD, double *aW,
                                            double *aB, double &aBn)
{
    const double d = W[DEN], p = W[PRS];
    const double id = 1.0 / d;
    const double b[3] = {B[0], B[1], B[2]};
    const double dPb = 0.5 * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    double t[3] = {b[0], b[1], b[2]};
    if (BG)
        for (int x = 0; x < 3; x++) t[x] += B0[x];
    const double tt = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    const double va2 = BG ? tt * id : 2 * dPb * id;
    const double den = 1 + va2 * m.iCa2;
    const double vac2 = BORIS ? va2 / den : va2;
    const double n[3] = {T[0], T[1], T[2]};
    const double vn = n[0] * W[1] + n[1] * W[2] + n[2] * W[3];
    const double lam = 1 / (2 * p * id + vac2);
    const double sl = sqrt(lam);
    const double z = s * sl * vn;
    double ez;
    const double vn0 = 0.5 * gam_erfc_exp(z, ez);
    const double bdb0 = BG ? B0[0] * b[0] + B0[1] * b[1] + B0[2] * b[2] : 0.0;
    double avn0 = 0.0, ndot = 0.0, bdot = 0.0;
    for (int x = 0; x < 3; x++) {
        avn0 += aF[x] * (dPb * n[x] - b[x] * Bn);
        if (BG) avn0 += aF[x] * (-(B0[x] * Bn) - b[x] * B0n + bdb0 * n[x]);
        ndot += aF[x] * n[x];
        bdot += aF[x] * b[x];
    }
    double adPb = vn0 * ndot;
    aBn += -vn0 * bdot;
    double ab[3];
    for (int x = 0; x < 3; x++) ab[x] = -vn0 * Bn * aF[x];
    if (BG) {
        double b0dot = 0.0;
        for (int x = 0; x < 3; x++) b0dot += B0[x] * aF[x];
        aBn += -vn0 * b0dot;
        for (int x = 0; x < 3; x++) ab[x] += -vn0 * B0n * aF[x] + vn0 * B0[x] * ndot;
    }
    // bbD += tt / 2 (with the background) or |b|^2 / 2
    double at[3] = {0.0, 0.0, 0.0};
    for (int x = 0; x < 3; x++) at[x] = abbD * t[x];
    // VaD += sqrt(vac2) / 2
    double avac2 = (vac2 > 0.0) ? 0.25 / sqrt(vac2) * aVaD : 0.0;
    const double az = -GAM_RSQRT_PI * ez * avn0;
    const double avn = s * sl * az;
    const double alam = (s * vn * az) * 0.5 / sl;
    const double aX = -lam * lam * alam;
    const double ap = 2 * id * aX;
    double aid = 2 * p * aX;
    avac2 += aX;
    const double ava2 = BORIS ? avac2 / (den * den) : avac2;
    if (BG) {
        const double att = id * ava2;
        aid += tt * ava2;
        for (int x = 0; x < 3; x++) at[x] += 2 * t[x] * att;
    } else {
        adPb += 2 * id * ava2;
        aid += 2 * dPb * ava2;
    }
    for (int x = 0; x < 3; x++) aB[x] += ab[x] + b[x] * adPb + at[x];
    aW[DEN] += -id * id * aid;
    aW[PRS] += ap;
    aW[1] += n[0] * avn;
    aW[2] += n[1] * avn;
    aW[3] += n[2] * avn;
}

// ---------------------------------------------------------- cell B
// Coefficients of gamCellBxyz (cell_b): dB_x / d(face flux) for the six
// faces of cell c in the order I, J, K (low), I+1, J+1, K+1 (high);
// af[f] += sum_x aB[x] dB_x/df.
__device__ __forceinline__ void cell_b_adj(const Geo &g, const double *met, long c, const double *aB,
                                           double *af)
{
    const long N3 = g.N3;
    const long fI = c, fJ = c, fK = c, fIp = c + 1, fJp = c + g.P, fKp = c + g.PP;
    const double idV = 1.0 / met[M_VOL * N3 + c];
    for (int x = 0; x < 3; x++) {
        const double xc = met[(M_XYZCC + x) * N3 + c];
        const double a = aB[x] * idV;
        af[0] += a * (-met[(M_XFC + x + 0) * N3 + fI] + xc);
        af[1] += a * (-met[(M_XFC + x + 3) * N3 + fJ] + xc);
        af[2] += a * (-met[(M_XFC + x + 6) * N3 + fK] + xc);
        af[3] += a * (met[(M_XFC + x + 0) * N3 + fIp] - xc);
        af[4] += a * (met[(M_XFC + x + 3) * N3 + fJp] - xc);
        af[5] += a * (met[(M_XFC + x + 6) * N3 + fKp] - xc);
    }
}

// ---------------------------------------------------------- update
// CellMaxwell (the Maxwell branch of k_update): U' = Maxwell(U, dM), with
// dt the step's dt.  aUp = adjoint of U'; adds into aU (5) and adM (3).
__device__ __forceinline__ void maxwell_adj(const MParm &m, const double *U, const double *dM, double dt,
                                            const double *aUp, double *aU, double *adM)
{
    const double d = dmax(U[DEN], m.dFloor);
    const double id = 1.0 / d;
    const double mo[3] = {U[MOMX], U[MOMY], U[MOMZ]};
    const double mp[3] = {mo[0] + dt * dM[0], mo[1] + dt * dM[1], mo[2] + dt * dM[2]};
    const double m2 = mo[0] * mo[0] + mo[1] * mo[1] + mo[2] * mo[2];
    const double mp2 = mp[0] * mp[0] + mp[1] * mp[1] + mp[2] * mp[2];
    const double pr = (m.gamma - 1) * (U[ENER] - 0.5 * m2 * id);
    double ad = aUp[DEN];
    double aid = 0.5 * mp2 * aUp[ENER];
    double amp[3], am[3];
    for (int x = 0; x < 3; x++) {
        amp[x] = aUp[1 + x] + mp[x] * id * aUp[ENER];
        am[x] = amp[x];
        adM[x] += dt * amp[x];
    }
    if (pr > m.pFloor) {
        const double apr = m.igm1 * aUp[ENER];
        aU[ENER] += (m.gamma - 1) * apr;
        const double ake = -(m.gamma - 1) * apr;
        for (int x = 0; x < 3; x++) am[x] += mo[x] * id * ake;
        aid += 0.5 * m2 * ake;
    }
    ad += -id * id * aid;
    if (U[DEN] > m.dFloor) aU[DEN] += ad;
    for (int x = 0; x < 3; x++) aU[1 + x] += am[x];
}

// ---------------------------------------------------------- predictor
// PredictCell (gas part of k_predict) at one cell: Q = State gas, oQ =
// oState gas, w = AB2 weight, vol = cell volume.  Adjoints of the
// predictor's outputs: aHG (half-state conserved, H_GAS), aWH (V * half
// primitives, H_WH), aWN (limiter primitives, H_WN).  Adds into aQ, aoQ,
// and, where aw is given, the adjoint of the AB2 weight w into *aw.
__device__ __forceinline__ void predict_cell_adj(const MParm &m, const double *Q, const double *oQ,
                                                 double w, double vol, const double *aHG,
                                                 const double *aWH, const double *aWN, double *aQ,
                                                 double *aoQ, double *aw = nullptr)
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
    if (aw)
        for (int v = 0; v < 5; v++) *aw += (W[v] - oW[v]) * apW[v];
    c2p_cell_adj(m, Q, aW, aQ);
    c2p_cell_adj(m, oQ, aoW, aoQ);
    if (m.doLFM) c2p_block_adj(m, Q, aWN, aQ);
}

// ---------------------------------------------------------- edge EMF
// Inputs of one edge EMF (k_emf) gathered into local arrays:
//   A1[8], V1[3][8]: areas and face velocities of the T1 faces along T2
//   mH1[8], mL1[4]: half-state and limiter-state fluxes of those faces
//   A2[8], mH2[8], mL2[4]: the same for the T2 faces along T1
//   te1[3], te2[3]: edge triad rows; tb[4] = xni, yni, xnj, ynj (Teb);
//   D4[4]: half-state densities of the four cells around the edge;
//   b01, b02: background edge field (0 without); edge length; cap = the
//   vDiff bound CFL * edge / dt.
// AV[d] = wsum8 over n of A1[n] V1[d][n] (V1 the face velocities of the
// Mom2Face average): the caller forms the sums as it loads V1, which never
// has to be held whole
struct EdgeIn {
    double A1[8], AV[3], mH1[8], mL1[4];
    double A2[8], mH2[8], mL2[4];
    double te1[3], te2[3], tb[4], D4[4];
    double b01, b02, edge, cap;
};
struct EdgeAdj {
    double V1[3][8], mH1[8], mL1[4], mH2[8], mL2[4], D4[4];
};

// forward (k_emf's arithmetic), also returning the pieces the adjoint needs
template <int RM, bool BORIS>
__device__ __forceinline__ double edge_emf(const MParm &m, const EdgeIn &in, double *l1, double *r1,
                                           double *l2, double *r2)
{
    const double rdA = 1.0 / dmax(wsum8([&](int n) { return in.A1[n]; }), GAM_TINY);
    double vc[3];
    for (int d = 0; d < 3; d++) vc[d] = in.AV[d] * rdA;
    const double v1 = vc[0] * in.te1[0] + vc[1] * in.te1[1] + vc[2] * in.te1[2];
    const double v2 = vc[0] * in.te2[0] + vc[1] * in.te2[1] + vc[2] * in.te2[2];
    double vDiff = sqrt(v1 * v1 + v2 * v2);
    const double Dc = 0.25 * (in.D4[0] + in.D4[1] + in.D4[2] + in.D4[3]);
    double q[8], nq[4];
    for (int n = 0; n < 8; n++) q[n] = in.mH1[n] / dmax(in.A1[n], GAM_TINY);
    for (int n = 0; n < 4; n++) nq[n] = in.mL1[n] / dmax(in.A1[n + 2], GAM_TINY);
    const double Vi1 = wsum8([&](int n) { return in.A1[n]; });
    recon_lr<RM>(m.pdmb, in.A1, 1.0 / Vi1, nq, q, *l1, *r1);
    for (int n = 0; n < 8; n++) q[n] = in.mH2[n] / dmax(in.A2[n], GAM_TINY);
    for (int n = 0; n < 4; n++) nq[n] = in.mL2[n] / dmax(in.A2[n + 2], GAM_TINY);
    const double Vi2 = wsum8([&](int n) { return in.A2[n]; });
    recon_lr<RM>(m.pdmb, in.A2, 1.0 / Vi2, nq, q, *l2, *r2);
    const double bT1 = 0.5 * (*l1 + *r1), bT2 = 0.5 * (*l2 + *r2);
    const double Jd = (*r2 - *l2) - (*r1 - *l1);
    const double xni = in.tb[0], yni = in.tb[1], xnj = in.tb[2], ynj = in.tb[3];
    const double detT = 1.0 / (xni * ynj - xnj * yni);
    const double b1 = detT * (ynj * bT1 - yni * bT2) + in.b01;
    const double b2 = detT * (-(xnj * bT1) + xni * bT2) + in.b02;
    double vA = sqrt(b1 * b1 + b2 * b2) / sqrt(Dc);
    if (BORIS) vA = m.Ca * vA / sqrt(m.Ca * m.Ca + vA * vA);
    vDiff = dmin(vDiff + vA, in.cap);
    return (-(v1 * b2 - v2 * b1) + m.Vd0 * vDiff * Jd) * in.edge;
}

// edge_emf's transpose up to its linear stencils (gathered by the cell and
// face side, k_emf_adj_gather): sv = the corner-velocity adjoint over the
// area sum (V1[d][n] gets W8(n) A1[n] sv[d]); sl, sr and anq the edge-B
// reconstructions' stencil scalars and limiter adjoints (recon_lr_w_adj_s)
// for the mH/A and mL/A values; aDc the corner density's adjoint (each D4
// gets aDc / 4).
// E is the edge's EMF (times the edge length, as k_emf stores it) and acap
// the adjoint of the vDiff cap (nonzero only where the cap binds).
struct EdgeAdjS {
    double sv[3], sl1, sr1, anq1[4], sl2, sr2, anq2[4], aDc;
    double E, acap;
};

template <int RM, bool BORIS>
__device__ __forceinline__ void edge_emf_adj_s(const MParm &m, const EdgeIn &in, double aE, EdgeAdjS &o)
{
    double l1, r1, l2, r2;
    o.E = edge_emf<RM, BORIS>(m, in, &l1, &r1, &l2, &r2);
    // forward pieces again
    const double rdA = 1.0 / dmax(wsum8([&](int n) { return in.A1[n]; }), GAM_TINY);
    double vc[3];
    for (int d = 0; d < 3; d++) vc[d] = in.AV[d] * rdA;
    const double v1 = vc[0] * in.te1[0] + vc[1] * in.te1[1] + vc[2] * in.te1[2];
    const double v2 = vc[0] * in.te2[0] + vc[1] * in.te2[1] + vc[2] * in.te2[2];
    const double vD = sqrt(v1 * v1 + v2 * v2);
    const double Dc = 0.25 * (in.D4[0] + in.D4[1] + in.D4[2] + in.D4[3]);
    const double bT1 = 0.5 * (l1 + r1), bT2 = 0.5 * (l2 + r2);
    const double Jd = (r2 - l2) - (r1 - l1);
    const double xni = in.tb[0], yni = in.tb[1], xnj = in.tb[2], ynj = in.tb[3];
    const double detT = 1.0 / (xni * ynj - xnj * yni);
    const double b1 = detT * (ynj * bT1 - yni * bT2) + in.b01;
    const double b2 = detT * (-(xnj * bT1) + xni * bT2) + in.b02;
    const double nb = sqrt(b1 * b1 + b2 * b2);
    const double u = nb / sqrt(Dc);
    double vA = u;
    if (BORIS) vA = m.Ca * u / sqrt(m.Ca * m.Ca + u * u);
    const double vsum = vD + vA;
    const double vDiff = dmin(vsum, in.cap);
    // adjoint
    const double ae = in.edge * aE;
    double av1 = -b2 * ae, av2 = b1 * ae, ab1 = v2 * ae, ab2 = -v1 * ae;
    const double avDiff = m.Vd0 * Jd * ae, aJd = m.Vd0 * vDiff * ae;
    const double avs = (vsum < in.cap) ? avDiff : 0.0;
    o.acap = (vsum < in.cap) ? 0.0 : avDiff;
    double au = avs;
    if (BORIS) {
        const double c2 = m.Ca * m.Ca + u * u;
        au = avs * m.Ca * m.Ca * m.Ca / (c2 * sqrt(c2));
    }
    const double anb = au / sqrt(Dc);
    const double aDc = -0.5 * u / Dc * au;
    if (nb > 0.0) { ab1 += b1 / nb * anb; ab2 += b2 / nb * anb; }
    const double abT1 = detT * (ynj * ab1 - xnj * ab2);
    const double abT2 = detT * (-yni * ab1 + xni * ab2);
    double al1 = aJd + 0.5 * abT1, ar1 = -aJd + 0.5 * abT1;
    double al2 = -aJd + 0.5 * abT2, ar2 = aJd + 0.5 * abT2;
    if (vD > 0.0) { av1 += v1 / vD * avs; av2 += v2 / vD * avs; }
    // edge B reconstructions
    {
        double qv[8], nq[4];
        for (int n = 0; n < 8; n++) qv[n] = in.A1[n] * (in.mH1[n] / dmax(in.A1[n], GAM_TINY));
        for (int n = 0; n < 4; n++) nq[n] = in.mL1[n] / dmax(in.A1[n + 2], GAM_TINY);
        const double Vi = wsum8([&](int n) { return in.A1[n]; });
        for (int n = 0; n < 4; n++) o.anq1[n] = 0.0;
        recon_lr_w_adj_s<RM>(m.pdmb, 1.0 / Vi, nq, qv, al1, ar1, o.anq1, o.sl1, o.sr1);
    }
    {
        double qv[8], nq[4];
        for (int n = 0; n < 8; n++) qv[n] = in.A2[n] * (in.mH2[n] / dmax(in.A2[n], GAM_TINY));
        for (int n = 0; n < 4; n++) nq[n] = in.mL2[n] / dmax(in.A2[n + 2], GAM_TINY);
        const double Vi = wsum8([&](int n) { return in.A2[n]; });
        for (int n = 0; n < 4; n++) o.anq2[n] = 0.0;
        recon_lr_w_adj_s<RM>(m.pdmb, 1.0 / Vi, nq, qv, al2, ar2, o.anq2, o.sl2, o.sr2);
    }
    // corner velocity and density
    for (int d = 0; d < 3; d++) o.sv[d] = rdA * (in.te1[d] * av1 + in.te2[d] * av2);
    o.aDc = aDc;
}

// the stencil weight of point n of the transposes split above
template <int RM>
__device__ __forceinline__ double stencil_adj(int n, double sl, double sr)
{
    if (RM == 1) return (n <= 6 ? UP7(n) * sl : 0.0) + (n >= 1 ? UP7(7 - n) * sr : 0.0);
    return W8(n) * sl;
}

template <int RM, bool BORIS>
__device__ __forceinline__ void edge_emf_adj(const MParm &m, const EdgeIn &in, double aE, EdgeAdj &out)
{
    EdgeAdjS t;
    edge_emf_adj_s<RM, BORIS>(m, in, aE, t);
    for (int n = 0; n < 8; n++) {
        out.mH1[n] += in.A1[n] * stencil_adj<RM>(n, t.sl1, t.sr1) / dmax(in.A1[n], GAM_TINY);
        out.mH2[n] += in.A2[n] * stencil_adj<RM>(n, t.sl2, t.sr2) / dmax(in.A2[n], GAM_TINY);
        for (int d = 0; d < 3; d++) out.V1[d][n] += W8(n) * in.A1[n] * t.sv[d];
    }
    for (int n = 0; n < 4; n++) {
        out.mL1[n] += t.anq1[n] / dmax(in.A1[n + 2], GAM_TINY);
        out.mL2[n] += t.anq2[n] / dmax(in.A2[n + 2], GAM_TINY);
        out.D4[n] += 0.25 * t.aDc;
    }
}

// ---------------------------------------------------------- CellDT
// cell_dt (CellDT, gridutils.F90:49-106) on conserved U and total cell B:
// adds adt * d(cell_dt)/dU into aU (5) and adt * d(cell_dt)/dB into aB (3).
template <bool BORIS>
__device__ __forceinline__ void cell_dt_adj(const MParm &m, const double *U, double bx, double by,
                                            double bz, double dl, double adt, double *aU, double *aB)
{
    const double rho = U[DEN];
    const double ir = 1.0 / rho;
    const double v[3] = {U[MOMX] * ir, U[MOMY] * ir, U[MOMZ] * ir};
    const double mv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double ke = 0.5 * rho * (mv * mv);
    const double p = (m.gamma - 1) * (U[ENER] - ke);
    const double cs = sqrt(m.gamma * p * ir);
    double u = 0.0, va = 0.0, nb = 0.0;
    if (m.doMHD) {
        nb = sqrt(bx * bx + by * by + bz * bz);
        u = nb / sqrt(rho);
        va = u;
        if (BORIS) va = m.Ca * u / sqrt(m.Ca * m.Ca + u * u);
    }
    const double f = sqrt(cs * cs + va * va);
    const double vcfl = mv + f + 0.0;
    const double dt = m.CFL * dl / vcfl;
    // reverse
    const double avc = -dt / vcfl * adt;
    double amv = avc, arho = 0.0, air = 0.0;
    const double af = avc;
    double acs = 0.0, ava = 0.0;
    if (f > 0.0) { acs = af * cs / f; ava = af * va / f; }
    if (m.doMHD) {
        double au = ava;
        if (BORIS) {
            const double c2 = m.Ca * m.Ca + u * u;
            au = ava * m.Ca * m.Ca * m.Ca / (c2 * sqrt(c2));
        }
        const double anb = au / sqrt(rho);
        arho += -0.5 * u / rho * au;
        if (nb > 0.0) {
            aB[0] += bx / nb * anb;
            aB[1] += by / nb * anb;
            aB[2] += bz / nb * anb;
        }
    }
    double ap = 0.0;
    if (cs > 0.0) {
        const double t = acs * m.gamma / (2.0 * cs);
        ap = t * ir;
        air += t * p;
    }
    aU[ENER] += (m.gamma - 1) * ap;
    const double ake = -(m.gamma - 1) * ap;
    arho += 0.5 * (mv * mv) * ake;
    amv += rho * mv * ake;
    if (mv > 0.0)
        for (int x = 0; x < 3; x++) {
            const double avx = v[x] / mv * amv;
            aU[1 + x] += ir * avx;
            air += U[1 + x] * avx;
        }
    arho += -ir * ir * air;
    aU[DEN] += arho;
}

// CellBoris (the Boris branch of k_update), single fluid.  Inputs: oU =
// State(n) gas, dH = gas flux divergence, dM = magnetic stress divergence
// (with dpB0), B = total cell B of the new state, oB = that of State(n),
// dtM = Model%dt (Reynolds), dt = the advance dt.  aUp = adjoint of the
// updated gas; adds into aoU, adH, adM, aB, aoB.
__device__ __forceinline__ void boris_adj(const MParm &m, const double *oU, const double *dH,
                                          const double *dM, const double *B, const double *oB, double dtM,
                                          double dt, const double *aUp, double *aoU, double *adH,
                                          double *adM, double *aB, double *aoB)
{
    double U[5], W2[5];
    for (int v = 0; v < 5; v++) U[v] = oU[v] + dtM * dH[v];
    c2p_cell(m, U, W2);
    const double D2 = dmax(W2[DEN], m.dFloor), P2 = dmax(W2[PRS], m.pFloor);
    const double D0 = dmax(oU[DEN], m.dFloor);
    const double iD0 = 1.0 / D0, iD2 = 1.0 / D2;
    double V0[3], B1[3], dPG[3], dPM[3];
    for (int x = 0; x < 3; x++) {
        V0[x] = oU[1 + x] * iD0;
        B1[x] = 0.5 * (B[x] + oB[x]);
        dPG[x] = dt * dH[1 + x];
        dPM[x] = dt * dM[x];
    }
    const double b2 = B1[0] * B1[0] + B1[1] * B1[1] + B1[2] * B1[2];
    const double alpha = b2 * (iD2 * m.iCa2);
    const double dRho = dt * dH[DEN];
    const double pCon = 1.0 / (1.0 + alpha);
    const bool bor = alpha > GAM_TINY;
    double bh[3] = {0.0, 0.0, 0.0}, sGM[3], dMom[3];
    const double nb = sqrt(b2);
    const bool nbOk = nb > GAM_TINY;
    if (nbOk)
        for (int x = 0; x < 3; x++) bh[x] = B1[x] / nb;
    for (int x = 0; x < 3; x++) sGM[x] = dPG[x] + dPM[x];
    const double pG = dPG[0] * bh[0] + dPG[1] * bh[1] + dPG[2] * bh[2];
    const double pS = sGM[0] * bh[0] + sGM[1] * bh[1] + sGM[2] * bh[2];
    const double pV = V0[0] * bh[0] + V0[1] * bh[1] + V0[2] * bh[2];
    for (int x = 0; x < 3; x++)
        dMom[x] = bor ? pG * bh[x] + pCon * (sGM[x] - pS * bh[x]) + alpha * pCon * dRho * (V0[x] - pV * bh[x])
                      : dPG[x] + dPM[x];
    double Wn[5];
    Wn[DEN] = D2;
    for (int x = 0; x < 3; x++) Wn[1 + x] = (D0 * V0[x] + dMom[x]) * iD2;
    Wn[PRS] = P2;
    // reverse
    double aWn[5] = {0, 0, 0, 0, 0};
    p2c_cell_adj(m, Wn, aUp, aWn);
    double aD2 = aWn[DEN], aP2 = aWn[PRS], aiD2 = 0.0, aD0 = 0.0;
    double aV0[3] = {0, 0, 0}, adMom[3], aDV[3];
    for (int x = 0; x < 3; x++) {
        aDV[x] = iD2 * aWn[1 + x];          // adjoint of D0 * V0[x]
        adMom[x] = iD2 * aWn[1 + x];
        aiD2 += (D0 * V0[x] + dMom[x]) * aWn[1 + x];
        aD0 += V0[x] * aDV[x];
        aV0[x] += D0 * aDV[x];
    }
    double adPG[3] = {0, 0, 0}, adPM[3] = {0, 0, 0}, aB1[3] = {0, 0, 0}, aalpha = 0.0, adRho = 0.0;
    if (bor) {
        const double K = alpha * pCon * dRho;
        double apG = 0.0, apS = 0.0, apV = 0.0, apCon = 0.0, abh[3] = {0, 0, 0}, asGM[3];
        for (int x = 0; x < 3; x++) {
            const double a = adMom[x];
            apG += bh[x] * a;
            abh[x] += (pG - pCon * pS - K * pV) * a;
            apCon += (sGM[x] - pS * bh[x]) * a + alpha * dRho * (V0[x] - pV * bh[x]) * a;
            asGM[x] = pCon * a;
            apS += -pCon * bh[x] * a;
            aalpha += pCon * dRho * (V0[x] - pV * bh[x]) * a;
            adRho += alpha * pCon * (V0[x] - pV * bh[x]) * a;
            aV0[x] += K * a;
            apV += -K * bh[x] * a;
        }
        for (int x = 0; x < 3; x++) {
            adPG[x] += bh[x] * apG;
            abh[x] += dPG[x] * apG;
            asGM[x] += bh[x] * apS;
            abh[x] += sGM[x] * apS;
            aV0[x] += bh[x] * apV;
            abh[x] += V0[x] * apV;
            adPG[x] += asGM[x];
            adPM[x] += asGM[x];
        }
        if (nbOk) {
            const double bda = bh[0] * abh[0] + bh[1] * abh[1] + bh[2] * abh[2];
            for (int x = 0; x < 3; x++) aB1[x] += (abh[x] - bh[x] * bda) / nb;
        }
        aalpha += -pCon * pCon * apCon;
    } else {
        for (int x = 0; x < 3; x++) { adPG[x] += adMom[x]; adPM[x] += adMom[x]; }
    }
    // alpha = |B1|^2 / (D2 Ca^2)
    const double ab2 = iD2 * m.iCa2 * aalpha;
    aiD2 += b2 * m.iCa2 * aalpha;
    for (int x = 0; x < 3; x++) {
        aB1[x] += 2 * B1[x] * ab2;
        aB[x] += 0.5 * aB1[x];
        aoB[x] += 0.5 * aB1[x];
        adM[x] += dt * adPM[x];
        adH[1 + x] += dt * adPG[x];
    }
    adH[DEN] += dt * adRho;
    // V0 = oU_mom / D0, D0 = max(oU_rho, dFloor)
    double aiD0 = 0.0;
    for (int x = 0; x < 3; x++) {
        aoU[1 + x] += iD0 * aV0[x];
        aiD0 += oU[1 + x] * aV0[x];
    }
    aD0 += -iD0 * iD0 * aiD0;
    if (oU[DEN] > m.dFloor) aoU[DEN] += aD0;
    aD2 += -iD2 * iD2 * aiD2;
    double aW2[5] = {0, 0, 0, 0, 0}, aU[5] = {0, 0, 0, 0, 0};
    if (W2[DEN] > m.dFloor) aW2[DEN] = aD2;
    if (W2[PRS] > m.pFloor) aW2[PRS] = aP2;
    c2p_cell_adj(m, U, aW2, aU);
    for (int v = 0; v < 5; v++) {
        aoU[v] += aU[v];
        adH[v] += dtM * aU[v];
    }
}
