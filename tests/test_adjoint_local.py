"""Reverse-mode products of the hydrodynamic per-point functions
(mage_amd/csrc/gam_adj.h) against central differences of the forward
functions they differentiate (gam_phys.h), both compiled for the host from
the same headers.

For a function y = f(x), random x, direction dx and output weights a, the
adjoint must satisfy a . (f(x + e dx) - f(x - e dx)) / (2 e) = (J^T a) . dx.
Smooth functions: agreement to 1e-7 of the sum of the absolute terms of
either product, with e = 1e-6 (central-difference truncation ~e^2,
rounding ~1e-16/e).  Piecewise-linear limiter functions:
exact within a branch, so e = 1e-9 and at most 1 % of samples may straddle
a branch switch; every other sample must agree to 1e-6.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = r'''
#include "gam_adj.h"
#include "gam_wind.h"
static MParm mp(double gamma, double dF, double pF, int lfm)
{
    MParm m = {};
    m.gamma = gamma; m.igm1 = 1.0 / (gamma - 1.0); m.dFloor = dF; m.pFloor = pF; m.doLFM = lfm;
    return m;
}
extern "C" {
void f_gas(double g, double s, const double *W, const double *T, double *y)
{ double dW[5] = {0, 0, 0, 0, 0}; for (int v = 0; v < 5; v++) y[v] = 0; gas_kin(1.0 / (g - 1.0), s, W, T, y, dW); }
void a_gas(double g, double s, const double *W, const double *T, const double *a, double *ax)
{ for (int v = 0; v < 5; v++) ax[v] = 0; gas_kin_adj(1.0 / (g - 1.0), s, W, T, a, ax); }
void f_c2p(double g, double dF, double pF, const double *U, double *y) { MParm m = mp(g, dF, pF, 1); c2p_cell(m, U, y); }
void a_c2p(double g, double dF, double pF, const double *U, const double *a, double *ax)
{ MParm m = mp(g, dF, pF, 1); for (int v = 0; v < 5; v++) ax[v] = 0; c2p_cell_adj(m, U, a, ax); }
void f_p2c(double g, double dF, double pF, const double *W, double *y) { MParm m = mp(g, dF, pF, 1); p2c_cell(m, W, y); }
void a_p2c(double g, double dF, double pF, const double *W, const double *a, double *ax)
{ MParm m = mp(g, dF, pF, 1); for (int v = 0; v < 5; v++) ax[v] = 0; p2c_cell_adj(m, W, a, ax); }
void f_blk(double g, double dF, double pF, const double *U, double *y) { MParm m = mp(g, dF, pF, 1); c2p_block(m, U, y); }
void a_blk(double g, double dF, double pF, const double *U, const double *a, double *ax)
{ MParm m = mp(g, dF, pF, 1); for (int v = 0; v < 5; v++) ax[v] = 0; c2p_block_adj(m, U, a, ax); }
void f_rec(int rm, double b, double rVi, const double *x, double *y)
{   // x = qv[8], nq[4]
    if (rm) recon_lr_w<1>(b, rVi, x + 8, x, y[0], y[1]); else recon_lr_w<0>(b, rVi, x + 8, x, y[0], y[1]);
}
void a_rec(int rm, double b, double rVi, const double *x, const double *a, double *ax)
{
    for (int n = 0; n < 12; n++) ax[n] = 0;
    if (rm) recon_lr_w_adj<1>(b, rVi, x + 8, x, a[0], a[1], ax + 8, ax);
    else recon_lr_w_adj<0>(b, rVi, x + 8, x, a[0], a[1], ax + 8, ax);
}
void f_gas2(double g, double s, const double *W, const double *T, double *y)
{ for (int v = 0; v < 10; v++) y[v] = 0; gas_kin(1.0 / (g - 1.0), s, W, T, y, y + 5); }
void a_gas2(double g, double s, const double *W, const double *T, const double *a, double *ax)
{ for (int v = 0; v < 5; v++) ax[v] = 0; gas_kin_adj(1.0 / (g - 1.0), s, W, T, a, ax, a + 5); }
void f_mag(int bg, int bo, double s, const double *x, const double *T, double *y)
{   // x = W[5], B[3], Bn; y = mF[3], VaD, bbD
    MParm m = mp(5.0 / 3.0, 1e-12, 1e-12, 1); m.iCa2 = 1.0 / 4.0;
    double B0[3] = {0.3, -0.2, 0.5}, bbD = 0.0, VaD = 0.0;
    for (int v = 0; v < 3; v++) y[v] = 0;
    if (bg && bo) mag_kin<true, true>(m, s, x, x + 5, x[8], T, B0, 0.4, y, bbD, VaD);
    else if (bg) mag_kin<true, false>(m, s, x, x + 5, x[8], T, B0, 0.4, y, bbD, VaD);
    else if (bo) mag_kin<false, true>(m, s, x, x + 5, x[8], T, B0, 0.4, y, bbD, VaD);
    else mag_kin<false, false>(m, s, x, x + 5, x[8], T, B0, 0.4, y, bbD, VaD);
    y[3] = VaD; y[4] = bbD;
}
void a_mag(int bg, int bo, double s, const double *x, const double *T, const double *a, double *ax)
{
    MParm m = mp(5.0 / 3.0, 1e-12, 1e-12, 1); m.iCa2 = 1.0 / 4.0;
    double B0[3] = {0.3, -0.2, 0.5};
    for (int v = 0; v < 9; v++) ax[v] = 0;
    double aBn = 0.0;
    if (bg && bo) mag_kin_adj<true, true>(m, s, x, x + 5, x[8], T, B0, 0.4, a, a[3], a[4], ax, ax + 5, aBn);
    else if (bg) mag_kin_adj<true, false>(m, s, x, x + 5, x[8], T, B0, 0.4, a, a[3], a[4], ax, ax + 5, aBn);
    else if (bo) mag_kin_adj<false, true>(m, s, x, x + 5, x[8], T, B0, 0.4, a, a[3], a[4], ax, ax + 5, aBn);
    else mag_kin_adj<false, false>(m, s, x, x + 5, x[8], T, B0, 0.4, a, a[3], a[4], ax, ax + 5, aBn);
    ax[8] = aBn;
}
void f_boris(double dtM, double dt, const double *x, double *y)
{   // x = oU[5], dH[5], dM[3], B[3], oB[3] (19): the Boris branch of k_update
    MParm m = mp(5.0 / 3.0, 1e-12, 1e-12, 1); m.iCa2 = 1.0 / 4.0;
    const double *oU = x, *dH = x + 5, *dM = x + 10, *B = x + 13, *oB = x + 16;
    double U[5], W2[5];
    for (int v = 0; v < 5; v++) U[v] = oU[v] + dtM * dH[v];
    c2p_cell(m, U, W2);
    const double D2 = dmax(W2[DEN], m.dFloor), P2 = dmax(W2[PRS], m.pFloor), D0 = dmax(oU[DEN], m.dFloor);
    const double iD0 = 1.0 / D0, iD2 = 1.0 / D2;
    double V0[3] = {oU[1] * iD0, oU[2] * iD0, oU[3] * iD0};
    double B1[3] = {0.5 * (B[0] + oB[0]), 0.5 * (B[1] + oB[1]), 0.5 * (B[2] + oB[2])};
    const double alpha = (B1[0] * B1[0] + B1[1] * B1[1] + B1[2] * B1[2]) * (iD2 * m.iCa2);
    const double dRho = dt * dH[DEN];
    double dPG[3] = {dt * dH[1], dt * dH[2], dt * dH[3]}, dPM[3] = {dt * dM[0], dt * dM[1], dt * dM[2]};
    const double pCon = 1.0 / (1.0 + alpha);
    double dMom[3];
    const double nb = sqrt(B1[0] * B1[0] + B1[1] * B1[1] + B1[2] * B1[2]);
    double bh[3] = {B1[0] / nb, B1[1] / nb, B1[2] / nb};
    double sGM[3] = {dPG[0] + dPM[0], dPG[1] + dPM[1], dPG[2] + dPM[2]};
    const double pG = dPG[0] * bh[0] + dPG[1] * bh[1] + dPG[2] * bh[2];
    const double pS = sGM[0] * bh[0] + sGM[1] * bh[1] + sGM[2] * bh[2];
    const double pV = V0[0] * bh[0] + V0[1] * bh[1] + V0[2] * bh[2];
    for (int x = 0; x < 3; x++)
        dMom[x] = pG * bh[x] + pCon * (sGM[x] - pS * bh[x]) + alpha * pCon * dRho * (V0[x] - pV * bh[x]);
    double Wn[5] = {D2, (D0 * V0[0] + dMom[0]) * iD2, (D0 * V0[1] + dMom[1]) * iD2, (D0 * V0[2] + dMom[2]) * iD2, P2};
    p2c_cell(m, Wn, y);
}
void a_boris(double dtM, double dt, const double *x, const double *a, double *ax)
{
    MParm m = mp(5.0 / 3.0, 1e-12, 1e-12, 1); m.iCa2 = 1.0 / 4.0;
    for (int v = 0; v < 19; v++) ax[v] = 0;
    boris_adj(m, x, x + 5, x + 10, x + 13, x + 16, dtM, dt, a, ax, ax + 5, ax + 10, ax + 13, ax + 16);
}
static Geo cb_geo() { Geo g = {}; g.P = 4; g.PP = 16; g.N3 = 64; return g; }
void f_cellb(const double *met, const double *f, double *y)
{
    Geo g = cb_geo();
    cell_b(g, met, 21, f[0], f[1], f[2], f[3], f[4], f[5], y);
}
void a_cellb(const double *met, const double *f, const double *a, double *ax)
{
    Geo g = cb_geo();
    for (int v = 0; v < 6; v++) ax[v] = 0;
    cell_b_adj(g, met, 21, a, ax);
}
void f_maxw(double g, double dt, const double *x, double *y)
{   // x = U[5], dM[3]: the Maxwell branch of k_update
    MParm m = mp(g, 1e-12, 1e-12, 1);
    const double *U = x, *dM = x + 5;
    const double d = dmax(U[DEN], m.dFloor);
    double mx = U[MOMX], my = U[MOMY], mz = U[MOMZ];
    const double id = 1.0 / d;
    double ke = 0.5 * (mx * mx + my * my + mz * mz) * id;
    const double p = dmax((m.gamma - 1) * (U[ENER] - ke), m.pFloor);
    mx = mx + dt * dM[0]; my = my + dt * dM[1]; mz = mz + dt * dM[2];
    ke = 0.5 * (mx * mx + my * my + mz * mz) * id;
    y[0] = d; y[1] = mx; y[2] = my; y[3] = mz; y[4] = ke + p * m.igm1;
}
void a_maxw(double g, double dt, const double *x, const double *a, double *ax)
{
    MParm m = mp(g, 1e-12, 1e-12, 1);
    for (int v = 0; v < 8; v++) ax[v] = 0;
    maxwell_adj(m, x, x + 5, dt, a, ax, ax + 5);
}
// edge EMF: x = V1[24], mH1[8], mL1[4], mH2[8], mL2[4], D4[4] (52); geometry fixed by geo[]
static void edge_fill(EdgeIn &in, const double *geo, const double *x)
{
    for (int n = 0; n < 8; n++) { in.A1[n] = geo[n]; in.A2[n] = geo[8 + n]; }
    for (int d = 0; d < 3; d++) { in.te1[d] = geo[16 + d]; in.te2[d] = geo[19 + d]; }
    for (int q = 0; q < 4; q++) in.tb[q] = geo[22 + q];
    in.b01 = geo[26]; in.b02 = geo[27]; in.edge = geo[28]; in.cap = geo[29];
    for (int d = 0; d < 3; d++) in.AV[d] = wsum8([&](int n) { return in.A1[n] * x[8 * d + n]; });
    for (int n = 0; n < 8; n++) { in.mH1[n] = x[24 + n]; in.mH2[n] = x[36 + n]; }
    for (int n = 0; n < 4; n++) { in.mL1[n] = x[32 + n]; in.mL2[n] = x[44 + n]; in.D4[n] = x[48 + n]; }
}
void f_emf(int rm, int boris, const double *geo, const double *x, double *y)
{
    MParm m = mp(5.0 / 3.0, 1e-12, 1e-12, 1); m.pdmb = 0.75; m.Vd0 = 0.5; m.Ca = 2.0;
    EdgeIn in; edge_fill(in, geo, x);
    double l1, r1, l2, r2;
    if (rm) y[0] = boris ? edge_emf<1, true>(m, in, &l1, &r1, &l2, &r2) : edge_emf<1, false>(m, in, &l1, &r1, &l2, &r2);
    else y[0] = boris ? edge_emf<0, true>(m, in, &l1, &r1, &l2, &r2) : edge_emf<0, false>(m, in, &l1, &r1, &l2, &r2);
}
void a_emf(int rm, int boris, const double *geo, const double *x, const double *a, double *ax)
{
    MParm m = mp(5.0 / 3.0, 1e-12, 1e-12, 1); m.pdmb = 0.75; m.Vd0 = 0.5; m.Ca = 2.0;
    EdgeIn in; edge_fill(in, geo, x);
    EdgeAdj o = {};
    if (rm) { if (boris) edge_emf_adj<1, true>(m, in, a[0], o); else edge_emf_adj<1, false>(m, in, a[0], o); }
    else { if (boris) edge_emf_adj<0, true>(m, in, a[0], o); else edge_emf_adj<0, false>(m, in, a[0], o); }
    for (int d = 0; d < 3; d++) for (int n = 0; n < 8; n++) ax[8 * d + n] = o.V1[d][n];
    for (int n = 0; n < 8; n++) { ax[24 + n] = o.mH1[n]; ax[36 + n] = o.mH2[n]; }
    for (int n = 0; n < 4; n++) { ax[32 + n] = o.mL1[n]; ax[44 + n] = o.mL2[n]; ax[48 + n] = o.D4[n]; }
}
void f_win(int boris, const double *nH, const double *x, double *y)
{   // x = U[5], B[3] of the last active cell; nH the boundary face normal
    MParm m = mp(5.0 / 3.0, 1e-12, 1e-12, 1); m.doBoris = boris; m.Ca = 0.8;
    wind_inner(m, x, x + 5, nH, y);
}
void a_win(int boris, const double *nH, const double *x, const double *a, double *ax)
{
    MParm m = mp(5.0 / 3.0, 1e-12, 1e-12, 1); m.doBoris = boris; m.Ca = 0.8;
    for (int v = 0; v < 8; v++) ax[v] = 0;
    wind_inner_adj(m, x, x + 5, nH, a, ax, ax + 5);
}
static WindDev wind2(double *tW, double *Q, double *B, const double *V)
{   // two equal samples with velocity V
    tW[0] = -1e3; tW[1] = 1e3;
    for (int i = 0; i < 2; i++) { Q[i] = 1.0; Q[2 + i] = V[0]; Q[4 + i] = V[1]; Q[6 + i] = V[2]; Q[8 + i] = 1e-3; }
    for (int i = 0; i < 6; i++) B[i] = 0.0;
    WindDev w = {tW, Q, B, 2, -1e3, 1e3, 2e3, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    return w;
}
void f_wgt(const double *pos, const double *x, double *y)
{   // x = the sample's velocity
    double tW[2], Q[10], B[6];
    WindDev w = wind2(tW, Q, B, x);
    const double vF[3] = {-0.5, 0.0, 0.0};
    y[0] = wind_wgt(w, vF, pos, 0.0);
}
void a_wgt(const double *pos, const double *x, const double *a, double *ax)
{
    double tW[2], Q[10], B[6], dw[3];
    WindDev w = wind2(tW, Q, B, x);
    const double vF[3] = {-0.5, 0.0, 0.0};
    wind_wgt_dv(w, vF, pos, 0.0, dw);
    for (int q = 0; q < 3; q++) ax[q] = a[0] * dw[q];
}
void f_pred(double g, int lfm, double w, double vol, const double *x, double *y)
{   // x = Q[5], oQ[5]; y = H_GAS 5, H_WH 5, H_WN 5 (as k_predict writes them)
    MParm m = mp(g, 1e-12, 1e-12, lfm);
    double W[5], oW[5], pW[5], pQ[5], Wb[5];
    c2p_cell(m, x + 5, oW); c2p_cell(m, x, W);
    for (int v = 0; v < 5; v++) pW[v] = ab2(W[v], oW[v], w);
    pW[DEN] = dmax(pW[DEN], m.dFloor); pW[PRS] = dmax(pW[PRS], m.pFloor);
    p2c_cell(m, pW, pQ);
    for (int v = 0; v < 5; v++) y[v] = pQ[v];
    c2p_block(m, pQ, Wb);
    for (int v = 0; v < 5; v++) y[5 + v] = vol * Wb[v];
    if (lfm) c2p_block(m, x, Wb);
    for (int v = 0; v < 5; v++) y[10 + v] = Wb[v];
}
void a_pred(double g, int lfm, double w, double vol, const double *x, const double *a, double *ax)
{
    MParm m = mp(g, 1e-12, 1e-12, lfm);
    for (int n = 0; n < 10; n++) ax[n] = 0;
    predict_cell_adj(m, x, x + 5, w, vol, a, a + 5, a + 10, ax, ax + 5);
}
// the predictor with its AB2 weight as an input: x = Q[5], oQ[5], w
void f_predw(double g, int lfm, double vol, const double *x, double *y) { f_pred(g, lfm, x[10], vol, x, y); }
void a_predw(double g, int lfm, double vol, const double *x, const double *a, double *ax)
{
    MParm m = mp(g, 1e-12, 1e-12, lfm);
    for (int n = 0; n < 11; n++) ax[n] = 0;
    predict_cell_adj(m, x, x + 5, x[10], vol, a, a + 5, a + 10, ax, ax + 5, ax + 10);
}
// CellDT: x = U[5], B[3] (total cell B), dl = 0.7
static MParm mcdt(int mhd)
{
    MParm m = mp(5.0 / 3.0, 1e-12, 1e-12, 1); m.doMHD = mhd; m.CFL = 0.3; m.Ca = 0.8;
    return m;
}
void f_cdt(int mhd, int boris, const double *x, double *y)
{
    MParm m = mcdt(mhd);
    y[0] = boris ? cell_dt<true>(m, x, x[5], x[6], x[7], 0.7) : cell_dt<false>(m, x, x[5], x[6], x[7], 0.7);
}
void a_cdt(int mhd, int boris, const double *x, const double *a, double *ax)
{
    MParm m = mcdt(mhd);
    for (int n = 0; n < 8; n++) ax[n] = 0;
    if (boris) cell_dt_adj<true>(m, x, x[5], x[6], x[7], 0.7, a[0], ax, ax + 5);
    else cell_dt_adj<false>(m, x, x[5], x[6], x[7], 0.7, a[0], ax, ax + 5);
}
// the edge EMF's value and dE/d(cap) (aE = 1) from its transpose
void a_emfcap(int rm, int boris, const double *geo, const double *x, double *out)
{
    MParm m = mp(5.0 / 3.0, 1e-12, 1e-12, 1); m.pdmb = 0.75; m.Vd0 = 0.5; m.Ca = 2.0;
    EdgeIn in; edge_fill(in, geo, x);
    EdgeAdjS o;
    if (rm) { if (boris) edge_emf_adj_s<1, true>(m, in, 1.0, o); else edge_emf_adj_s<1, false>(m, in, 1.0, o); }
    else { if (boris) edge_emf_adj_s<0, true>(m, in, 1.0, o); else edge_emf_adj_s<0, false>(m, in, 1.0, o); }
    out[0] = o.E; out[1] = o.acap;
}
// the update's dt share as the kernels form it: (dH . adH + dM . adM) / dt
// from the transposes, for a cell updated with dtM = dt = t.
// Maxwell: x = oU[5], dH[5], dM[3]; U' = Maxwell(oU + t dH, dM, t)
void f_maxt(double g, double t, const double *x, double *y)
{
    double u[8];
    for (int v = 0; v < 5; v++) u[v] = x[v] + t * x[5 + v];
    for (int q = 0; q < 3; q++) u[5 + q] = x[10 + q];
    f_maxw(g, t, u, y);
}
double d_maxt(double g, double t, const double *x, const double *a)
{
    double u[8], au[8];
    for (int v = 0; v < 5; v++) u[v] = x[v] + t * x[5 + v];
    for (int q = 0; q < 3; q++) u[5 + q] = x[10 + q];
    a_maxw(g, t, u, a, au);
    double s = 0.0;
    for (int v = 0; v < 5; v++) s += x[5 + v] * (t * au[v]);   // adH = dtM aU
    for (int q = 0; q < 3; q++) s += x[10 + q] * au[5 + q];
    return s / t;
}
// Boris: x as f_boris
void f_bort(double t, const double *x, double *y) { f_boris(t, t, x, y); }
double d_bort(double t, const double *x, const double *a)
{
    double ax[19];
    a_boris(t, t, x, a, ax);
    double s = 0.0;
    for (int v = 0; v < 5; v++) s += x[5 + v] * ax[5 + v];
    for (int q = 0; q < 3; q++) s += x[10 + q] * ax[10 + q];
    return s / t;
}
}
'''


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("adjloc")
    src = d / "adj_host.cpp"
    src.write_text(SRC)
    so = d / "libadj_host.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-std=c++17",
                    "-I", os.path.join(ROOT, "tests", "host_shim"),
                    "-I", os.path.join(ROOT, "mage_amd", "csrc"), str(src), "-o", str(so)], check=True)
    return C.CDLL(str(so))


DP = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(DP)


def _check(fwd, adj, x0, nout, rng, eps, smooth, nsamp=400, perturb=None):
    bad = 0
    worst = 0.0
    for _ in range(nsamp):
        x = perturb(rng) if perturb else x0(rng)
        dx = rng.standard_normal(x.size) * np.maximum(np.abs(x), 1e-3)
        a = rng.standard_normal(nout)
        yp = np.zeros(nout); ym = np.zeros(nout); ax = np.zeros(x.size)
        fwd(x + eps * dx, yp)
        fwd(x - eps * dx, ym)
        lhs = a @ (yp - ym) / (2 * eps)
        adj(x, a, ax)
        rhs = ax @ dx
        # scale free of cancellation between the terms of the two dot products
        scale = max(np.sum(np.abs(a * (yp - ym))) / (2 * eps), np.sum(np.abs(ax * dx)), 1e-300)
        err = abs(lhs - rhs) / scale
        if smooth:
            worst = max(worst, err)
        elif err > 1e-6:
            bad += 1
    if smooth:
        assert worst <= 1e-7, worst
    else:
        assert bad <= nsamp // 100, bad


def _triad(rng):
    n = rng.standard_normal(3); n /= np.linalg.norm(n)
    t1 = np.cross(n, rng.standard_normal(3)); t1 /= np.linalg.norm(t1)
    t2 = np.cross(n, t1)
    return np.concatenate([n, t1, t2])


@pytest.mark.parametrize("s", [-1.0, 1.0])
def test_gas_kinetic_flux_adjoint(lib, s):
    rng = np.random.default_rng(11)
    T = _triad(rng)
    g = 5.0 / 3.0

    def x0(r):
        return np.array([r.uniform(0.2, 2.0), r.standard_normal(), r.standard_normal(),
                         r.standard_normal(), r.uniform(0.05, 2.0)])
    W = None
    fwd = lambda x, y: lib.f_gas(C.c_double(g), C.c_double(s), _p(np.array([x[0], x[1], x[2], x[3], x[4]])), _p(T), _p(y))
    adj = lambda x, a, ax: lib.a_gas(C.c_double(g), C.c_double(s), _p(x), _p(T), _p(a), _p(ax))
    _check(fwd, adj, x0, 5, rng, 1e-6, True)


@pytest.mark.parametrize("name", ["c2p", "p2c", "blk"])
def test_conversion_adjoints(lib, name):
    rng = np.random.default_rng(12)
    g = 5.0 / 3.0
    f = getattr(lib, "f_" + name)
    a_ = getattr(lib, "a_" + name)

    def x0(r):
        d = r.uniform(0.3, 2.0)
        v = r.standard_normal(3) * 0.5
        if name == "p2c":
            return np.array([d, v[0], v[1], v[2], r.uniform(0.1, 2.0)])
        return np.array([d, d * v[0], d * v[1], d * v[2], 0.5 * d * v @ v + r.uniform(0.1, 2.0) * 1.5])
    fwd = lambda x, y: f(C.c_double(g), C.c_double(1e-12), C.c_double(1e-12), _p(x), _p(y))
    adj = lambda x, a, ax: a_(C.c_double(g), C.c_double(1e-12), C.c_double(1e-12), _p(x), _p(a), _p(ax))
    _check(fwd, adj, x0, 5, rng, 1e-6, True)


@pytest.mark.parametrize("rm", [0, 1])
def test_reconstruction_and_pdm_adjoint(lib, rm):
    rng = np.random.default_rng(13 + rm)
    vol = rng.uniform(0.8, 1.2, 8)
    Vi = sum(w * v for w, v in zip([-3, 29, -139, 533, 533, -139, 29, -3], vol)) / 840.0

    def x0(r):
        q = np.cumsum(r.standard_normal(8)) * 0.3 + r.standard_normal() * 0.2
        nq = q[2:6] + 0.05 * r.standard_normal(4)
        return np.concatenate([vol * q, nq])
    fwd = lambda x, y: lib.f_rec(rm, C.c_double(0.75), C.c_double(1.0 / Vi), _p(x), _p(y))
    adj = lambda x, a, ax: lib.a_rec(rm, C.c_double(0.75), C.c_double(1.0 / Vi), _p(x), _p(a), _p(ax))
    _check(fwd, adj, x0, 2, rng, 1e-9, False, nsamp=2000)


@pytest.mark.parametrize("lfm", [0, 1])
def test_predictor_cell_adjoint(lib, lfm):
    rng = np.random.default_rng(17)
    g = 5.0 / 3.0

    def cons(r):
        d = r.uniform(0.3, 2.0)
        v = r.standard_normal(3) * 0.5
        return np.array([d, d * v[0], d * v[1], d * v[2], 0.5 * d * v @ v + r.uniform(0.2, 2.0) * 1.5])

    def x0(r):
        Q = cons(r)
        return np.concatenate([Q, Q * (1 + 0.05 * r.standard_normal(5))])
    fwd = lambda x, y: lib.f_pred(C.c_double(g), lfm, C.c_double(0.5), C.c_double(1.3), _p(x), _p(y))
    adj = lambda x, a, ax: lib.a_pred(C.c_double(g), lfm, C.c_double(0.5), C.c_double(1.3), _p(x), _p(a), _p(ax))
    _check(fwd, adj, x0, 15, rng, 1e-6, True)


@pytest.mark.parametrize("s", [-1.0, 1.0])
def test_gas_flux_with_hogs_difference_adjoint(lib, s):
    rng = np.random.default_rng(21)
    T = _triad(rng)
    g = 5.0 / 3.0

    def x0(r):
        return np.array([r.uniform(0.2, 2.0), r.standard_normal(), r.standard_normal(),
                         r.standard_normal(), r.uniform(0.05, 2.0)])
    fwd = lambda x, y: lib.f_gas2(C.c_double(g), C.c_double(s), _p(x), _p(T), _p(y))
    adj = lambda x, a, ax: lib.a_gas2(C.c_double(g), C.c_double(s), _p(x), _p(T), _p(a), _p(ax))
    _check(fwd, adj, x0, 10, rng, 1e-6, True)


@pytest.mark.parametrize("s", [-1.0, 1.0])
@pytest.mark.parametrize("bg,bo", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_magnetic_flux_adjoint(lib, s, bg, bo):
    rng = np.random.default_rng(23 + bg + 2 * bo)
    T = _triad(rng)

    def x0(r):
        return np.concatenate([[r.uniform(0.3, 2.0)], r.standard_normal(3) * 0.5, [r.uniform(0.1, 2.0)],
                               r.standard_normal(3), [r.standard_normal()]])
    fwd = lambda x, y: lib.f_mag(bg, bo, C.c_double(s), _p(x), _p(T), _p(y))
    adj = lambda x, a, ax: lib.a_mag(bg, bo, C.c_double(s), _p(x), _p(T), _p(a), _p(ax))
    _check(fwd, adj, x0, 5, rng, 1e-6, True)


def test_boris_update_adjoint(lib):
    rng = np.random.default_rng(31)

    def x0(r):
        d = r.uniform(0.5, 2.0)
        v = r.standard_normal(3) * 0.3
        oU = [d, d * v[0], d * v[1], d * v[2], 0.5 * d * v @ v + r.uniform(0.5, 2.0) * 1.5]
        return np.concatenate([oU, r.standard_normal(5) * 0.1, r.standard_normal(3) * 0.2,
                               r.standard_normal(3), r.standard_normal(3)])
    fwd = lambda x, y: lib.f_boris(C.c_double(0.05), C.c_double(0.05), _p(x), _p(y))
    adj = lambda x, a, ax: lib.a_boris(C.c_double(0.05), C.c_double(0.05), _p(x), _p(a), _p(ax))
    _check(fwd, adj, x0, 5, rng, 1e-6, True)


def test_cell_b_adjoint(lib):
    rng = np.random.default_rng(25)
    met = rng.uniform(0.5, 1.5, 91 * 64)
    fwd = lambda x, y: lib.f_cellb(_p(met), _p(x), _p(y))
    adj = lambda x, a, ax: lib.a_cellb(_p(met), _p(x), _p(a), _p(ax))
    _check(fwd, adj, lambda r: r.standard_normal(6), 3, rng, 1e-6, True)


def test_maxwell_update_adjoint(lib):
    rng = np.random.default_rng(27)
    g = 5.0 / 3.0

    def x0(r):
        d = r.uniform(0.3, 2.0)
        v = r.standard_normal(3) * 0.5
        U = [d, d * v[0], d * v[1], d * v[2], 0.5 * d * v @ v + r.uniform(0.2, 2.0) * 1.5]
        return np.concatenate([U, r.standard_normal(3)])
    fwd = lambda x, y: lib.f_maxw(C.c_double(g), C.c_double(0.07), _p(x), _p(y))
    adj = lambda x, a, ax: lib.a_maxw(C.c_double(g), C.c_double(0.07), _p(x), _p(a), _p(ax))
    _check(fwd, adj, x0, 5, rng, 1e-6, True)


@pytest.mark.parametrize("rm,boris,cap", [(1, 0, 1e9), (0, 1, 1e9), (1, 1, 0.05)])
def test_edge_emf_adjoint(lib, rm, boris, cap):
    """Edge EMF (corner velocity, limited edge B, vDiff with Boris vA and
    the CFL cap) -- piecewise smooth through the limiter, like the
    reconstruction test."""
    rng = np.random.default_rng(29 + rm + 2 * boris)
    t1 = _triad(rng)
    geo = np.concatenate([rng.uniform(0.8, 1.2, 16), t1[3:6], t1[6:9], [1.0, 0.1, 0.05, 0.9],
                          [0.05, -0.02, 0.3, cap]])

    def x0(r):
        V = r.standard_normal(24) * 0.3
        b1 = np.cumsum(r.standard_normal(8)) * 0.2 + 1.0
        b2 = np.cumsum(r.standard_normal(8)) * 0.2 - 0.5
        return np.concatenate([V, b1, b1[2:6] + 0.02 * r.standard_normal(4), b2,
                               b2[2:6] + 0.02 * r.standard_normal(4), r.uniform(0.5, 1.5, 4)])
    fwd = lambda x, y: lib.f_emf(rm, boris, _p(geo), _p(x), _p(y))
    adj = lambda x, a, ax: lib.a_emf(rm, boris, _p(geo), _p(x), _p(a), _p(ax))
    _check(fwd, adj, x0, 1, rng, 1e-9, False, nsamp=2000)


@pytest.mark.parametrize("boris", [0, 1])
@pytest.mark.parametrize("fix", [True, False])
def test_inflow_interior_cell_adjoint(lib, boris, fix):
    """The inflow BC's interior primitives: CellC2P of the last active cell
    with the outflow velocity raised to the fast speed (Con2Fast sound and
    Boris-limited Alfven speeds) where it is below it (fix), or left as it
    is (a supersonic outflow, no fix)."""
    rng = np.random.default_rng(41 + boris + 2 * fix)
    nH = _triad(rng)[:3].copy()

    def x0(r):
        d = r.uniform(0.3, 2.0)
        v = r.standard_normal(3) * 0.3
        v += ((0.0 if fix else 8.0) - v @ nH) * nH + r.uniform(-0.2, 0.2) * nH
        U = [d, d * v[0], d * v[1], d * v[2], 0.5 * d * v @ v + r.uniform(0.1, 2.0) * 1.5]
        return np.concatenate([U, r.standard_normal(3)])
    fwd = lambda x, y: lib.f_win(boris, _p(nH), _p(x), _p(y))
    adj = lambda x, a, ax: lib.a_win(boris, _p(nH), _p(x), _p(a), _p(ax))
    _check(fwd, adj, x0, 5, rng, 1e-6, True)


def test_inflow_weight_direction_adjoint(lib):
    """The inflow blend weight (RampDown of the angle between the position
    and the upstream direction) against the sample's velocity, at positions
    on the downstream side where the weight falls from 1 to 0."""
    rng = np.random.default_rng(43)
    for _ in range(20):
        ph = rng.uniform(0.55, 0.95) * np.pi
        th = rng.uniform(0.2, 0.8) * np.pi
        pos = 20.0 * np.array([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)])
        fwd = lambda x, y: lib.f_wgt(_p(pos), _p(x), _p(y))
        adj = lambda x, a, ax: lib.a_wgt(_p(pos), _p(x), _p(a), _p(ax))
        _check(fwd, adj, lambda r: np.array([-0.5, 0.0, 0.0]) + r.standard_normal(3) * 0.1, 1, rng, 1e-6,
               True, nsamp=20)
        y = np.zeros(1)
        fwd(np.array([-0.5, 0.0, 0.0]), y)
        assert 0.0 <= y[0] <= 1.0


@pytest.mark.parametrize("lfm", [0, 1])
def test_predictor_weight_adjoint(lib, lfm):
    """The AB2 weight's adjoint (the predictor's share of dJ/d(dt))."""
    rng = np.random.default_rng(47 + lfm)
    g = 5.0 / 3.0

    def x0(r):
        d = r.uniform(0.3, 2.0)
        v = r.standard_normal(3) * 0.5
        Q = np.array([d, d * v[0], d * v[1], d * v[2], 0.5 * d * v @ v + r.uniform(0.2, 2.0) * 1.5])
        return np.concatenate([Q, Q * (1 + 0.05 * r.standard_normal(5)), [r.uniform(0.2, 0.8)]])
    fwd = lambda x, y: lib.f_predw(C.c_double(g), lfm, C.c_double(1.3), _p(x), _p(y))
    adj = lambda x, a, ax: lib.a_predw(C.c_double(g), lfm, C.c_double(1.3), _p(x), _p(a), _p(ax))
    _check(fwd, adj, x0, 15, rng, 1e-6, True)


@pytest.mark.parametrize("mhd,boris", [(0, 0), (1, 0), (1, 1)])
def test_cell_dt_adjoint(lib, mhd, boris):
    """CellDT (sound, Alfven with the Boris limit, flow speed) against its
    transpose: the CalcDT share of the dt-following gradient."""
    rng = np.random.default_rng(53 + mhd + 2 * boris)

    def x0(r):
        d = r.uniform(0.3, 2.0)
        v = r.standard_normal(3) * 0.5
        U = [d, d * v[0], d * v[1], d * v[2], 0.5 * d * v @ v + r.uniform(0.1, 2.0) * 1.5]
        return np.concatenate([U, r.standard_normal(3)])
    fwd = lambda x, y: lib.f_cdt(mhd, boris, _p(x), _p(y))
    adj = lambda x, a, ax: lib.a_cdt(mhd, boris, _p(x), _p(a), _p(ax))
    _check(fwd, adj, x0, 1, rng, 1e-6, True)


@pytest.mark.parametrize("rm,boris", [(1, 0), (0, 1)])
def test_edge_emf_cap_derivative(lib, rm, boris):
    """The EMF's value as its transpose recomputes it, and dE/d(cap) where
    the vDiff cap CFL edge / dt binds (0 where it does not)."""
    rng = np.random.default_rng(59 + rm)
    t1 = _triad(rng)
    lib.f_emf.argtypes = None
    nbind = 0
    for q in range(200):
        cap = rng.uniform(0.02, 3.0)
        geo = np.concatenate([rng.uniform(0.8, 1.2, 16), t1[3:6], t1[6:9], [1.0, 0.1, 0.05, 0.9],
                              [0.05, -0.02, 0.3, cap]])
        V = rng.standard_normal(24) * 0.3
        b1 = np.cumsum(rng.standard_normal(8)) * 0.2 + 1.0
        b2 = np.cumsum(rng.standard_normal(8)) * 0.2 - 0.5
        x = np.concatenate([V, b1, b1[2:6] + 0.02 * rng.standard_normal(4), b2,
                            b2[2:6] + 0.02 * rng.standard_normal(4), rng.uniform(0.5, 1.5, 4)])
        out = np.zeros(2)
        lib.a_emfcap(rm, boris, _p(geo), _p(x), _p(out))
        y = np.zeros(1)
        lib.f_emf(rm, boris, _p(geo), _p(x), _p(y))
        assert out[0] == y[0]
        e = 1e-7 * cap
        gp, gm = geo.copy(), geo.copy()
        gp[29] += e
        gm[29] -= e
        yp, ym = np.zeros(1), np.zeros(1)
        lib.f_emf(rm, boris, _p(gp), _p(x), _p(yp))
        lib.f_emf(rm, boris, _p(gm), _p(x), _p(ym))
        fd = (yp[0] - ym[0]) / (2 * e)
        if out[1] != 0.0:
            nbind += 1
        # central-difference rounding: ~1e-16 |E| / e
        tol = 1e-6 * max(abs(fd), abs(out[1])) + 4e-16 * max(abs(y[0]), 1.0) / e
        assert abs(fd - out[1]) <= tol, (q, fd, out[1])
    assert 20 <= nbind <= 180, nbind   # both sides of the cap were sampled


@pytest.mark.parametrize("kind", ["maxwell", "boris"])
def test_update_dt_share(lib, kind):
    """The update transposes' dt share, (dH . adH + dM . adM) / dt, equals
    the derivative of the updated gas along dt (the update depends on dt
    only through dt dH and dt dM)."""
    rng = np.random.default_rng(61)
    lib.d_maxt.restype = C.c_double
    lib.d_bort.restype = C.c_double
    for q in range(100):
        d = rng.uniform(0.5, 2.0)
        v = rng.standard_normal(3) * 0.3
        oU = [d, d * v[0], d * v[1], d * v[2], 0.5 * d * v @ v + rng.uniform(0.5, 2.0) * 1.5]
        x = np.concatenate([oU, rng.standard_normal(5) * 0.1, rng.standard_normal(3) * 0.2,
                            rng.standard_normal(3), rng.standard_normal(3)])
        a = rng.standard_normal(5)
        t = rng.uniform(0.01, 0.1)
        e = 1e-6 * t
        yp, ym = np.zeros(5), np.zeros(5)
        if kind == "maxwell":
            lib.f_maxt(C.c_double(5.0 / 3.0), C.c_double(t + e), _p(x), _p(yp))
            lib.f_maxt(C.c_double(5.0 / 3.0), C.c_double(t - e), _p(x), _p(ym))
            ad = lib.d_maxt(C.c_double(5.0 / 3.0), C.c_double(t), _p(x), _p(a))
        else:
            lib.f_bort(C.c_double(t + e), _p(x), _p(yp))
            lib.f_bort(C.c_double(t - e), _p(x), _p(ym))
            ad = lib.d_bort(C.c_double(t), _p(x), _p(a))
        fd = a @ (yp - ym) / (2 * e)
        scale = np.sum(np.abs(a * (yp - ym))) / (2 * e)
        assert abs(fd - ad) <= 1e-7 * max(scale, 1e-300), (q, fd, ad)


def test_adjoint_catches_a_planted_sign_error(lib):
    """The check fails for a wrong product: flip the sign of one adjoint
    component of the gas flux and the comparison must reject it."""
    rng = np.random.default_rng(19)
    T = _triad(rng)
    g = 5.0 / 3.0
    x = np.array([1.1, 0.3, -0.2, 0.4, 0.7])
    dx = rng.standard_normal(5)
    a = rng.standard_normal(5)
    e = 1e-6
    yp = np.zeros(5); ym = np.zeros(5); ax = np.zeros(5)
    lib.f_gas(C.c_double(g), C.c_double(1.0), _p(x + e * dx), _p(T), _p(yp))
    lib.f_gas(C.c_double(g), C.c_double(1.0), _p(x - e * dx), _p(T), _p(ym))
    lib.a_gas(C.c_double(g), C.c_double(1.0), _p(x), _p(T), _p(a), _p(ax))
    ax[4] = -ax[4]
    lhs = a @ (yp - ym) / (2 * e)
    assert abs(lhs - ax @ dx) > 1e-4 * abs(lhs)
