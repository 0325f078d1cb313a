// gam_wind.h -- the inflow series lookups of the wind boundary (device).
#pragma once
#include "gam_phys.h"


// the sample InterpWind takes at time t (discrete walls)
__device__ inline int wind_index(const WindDev &w, double t)
{
    int i0;
    if (t >= w.tMax) i0 = w.n - 1;
    else if (t <= w.tMin) i0 = 0;
    else {
        i0 = (int)floor((t - w.tMin) / w.dtW);
        if (i0 + 1 > w.n - 1) i0 = w.n - 2;
    }
    return i0;
}

__device__ inline void wind_interp(const WindDev &w, double t, double *rho, double *pr, double *V, double *B)
{   // InterpWind/InterpWgts with discrete walls (wind.F90:766-868)
    const int i0 = wind_index(w, t);
    *rho = w.Q[i0];
    *pr = w.Q[i0 + 4 * w.n];
    V[0] = w.Q[i0 + 1 * w.n]; V[1] = w.Q[i0 + 2 * w.n]; V[2] = w.Q[i0 + 3 * w.n];
    B[1] = w.B[i0 + w.n]; B[2] = w.B[i0 + 2 * w.n];
    B[0] = w.Bx0 + w.ByC * B[1] + w.BzC * B[2];
    if (t <= w.tMin) { B[0] = 0.0; B[1] = 0.0; B[2] = 0.0; }
}

__device__ inline void wind_front(const WindDev &w, double t, double *vF)
{   // RefreshWind (wind.F90:114-155)
    double r, p, V[3], B[3];
    wind_interp(w, t, &r, &p, V, B);
    const double vs = -sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
    const double bcc = sqrt(w.ByC * w.ByC + w.BzC * w.BzC);
    const double btt = sqrt(1 + bcc * bcc);
    if (bcc > GAM_TINY) {
        const double cB = w.ByC / bcc, sB = w.BzC / bcc, cF = 1.0 / btt, sF = bcc / btt;
        vF[0] = vs * (cF * cF);
        vF[1] = -vs * cF * sF * cB;
        vF[2] = -vs * cF * sF * sB;
    } else { vF[0] = vs; vF[1] = 0.0; vF[2] = 0.0; }
}

__device__ inline void wind_at(const WindDev &w, const double *vF, const double *x, double t,
                        double *rho, double *pr, double *V, double *B)
{   // GetWindAt (wind.F90:157-175)
    const double vm = sqrt(vF[0] * vF[0] + vF[1] * vF[1] + vF[2] * vF[2]);
    const double dl = ((x[0] - w.xW) * vF[0] + (x[1] - w.yW) * vF[1] + (x[2] - w.zW) * vF[2]) / (vm * vm);
    wind_interp(w, t - dl, rho, pr, V, B);
}

// wgtWind (wind.F90:180-203) with RampDown(dTh,90,90) (math.F90:97-111)
// for the sample velocity V that wind_at gives at x
__device__ inline double wind_wgt_v(const double *x, const double *V)
{
    double nh[3], vh[3];
    double na = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    if (na <= GAM_TINY) { nh[0] = nh[1] = nh[2] = 0.0; } else { nh[0] = x[0] / na; nh[1] = x[1] / na; nh[2] = x[2] / na; }
    double nv = sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
    if (nv <= GAM_TINY) { vh[0] = vh[1] = vh[2] = 0.0; } else { vh[0] = V[0] / nv; vh[1] = V[1] / nv; vh[2] = V[2] / nv; }
    const double dTh = acos(-vh[0] * nh[0] + -vh[1] * nh[1] + -vh[2] * nh[2]) * 180.0 / GAM_PI;
    if (dTh <= 90.0) return 1.0;
    if (dTh >= 180.0) return 0.0;
    const double rs = (dTh - 90.0) / 90.0;
    const double ea = -1.0 / (1.0 - rs * rs);
    return exp(ea + 1.0);
}

__device__ inline double wind_wgt(const WindDev &w, const double *vF, const double *x, double t)
{
    double r, p, V[3], B[3];
    wind_at(w, vF, x, t, &r, &p, V, B);
    return wind_wgt_v(x, V);
}


// time at which wind_at samples the series for position x (the lag of
// GetWindAt, wind.F90:157-175)
__device__ inline double wind_lag_time(const WindDev &w, const double *vF, const double *x, double t)
{
    const double vm = sqrt(vF[0] * vF[0] + vF[1] * vF[1] + vF[2] * vF[2]);
    const double dl = ((x[0] - w.xW) * vF[0] + (x[1] - w.yW) * vF[1] + (x[2] - w.zW) * vF[2]) / (vm * vm);
    return t - dl;
}

// d wind_wgt / d V of the sample it reads (the weight depends on the wind
// direction through RampDown of the angle to the position)
__device__ inline void wind_wgt_dv(const WindDev &w, const double *vF, const double *x, double t, double *dw)
{
    dw[0] = dw[1] = dw[2] = 0.0;
    double nh[3], r, p, V[3], B[3];
    const double na = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    if (na <= GAM_TINY) return;
    for (int q = 0; q < 3; q++) nh[q] = x[q] / na;
    wind_at(w, vF, x, t, &r, &p, V, B);
    const double nv = sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
    if (nv <= GAM_TINY) return;
    double vh[3];
    for (int q = 0; q < 3; q++) vh[q] = V[q] / nv;
    const double c = -(vh[0] * nh[0] + vh[1] * nh[1] + vh[2] * nh[2]);
    const double dTh = acos(c) * 180.0 / GAM_PI;
    if (dTh <= 90.0 || dTh >= 180.0) return;
    const double rs = (dTh - 90.0) / 90.0;
    const double q1 = 1.0 - rs * rs;
    const double wgt = exp(-1.0 / q1 + 1.0);
    const double dw_dTh = wgt * (-2.0 * rs / (q1 * q1)) / 90.0;
    const double dTh_dc = -(180.0 / GAM_PI) / sqrt(1.0 - c * c);
    const double vn = vh[0] * nh[0] + vh[1] * nh[1] + vh[2] * nh[2];
    for (int q = 0; q < 3; q++) dw[q] = dw_dTh * dTh_dc * (-(nh[q] - vh[q] * vn) / nv);
}
