// gam_common.h -- shared device-side definitions for the gfx950 GAMERA step.
//
// Device layout: every field is a structure-of-arrays "plane" of N3 doubles
// over ONE padded index space shared by cell-, face- and edge-centred data:
//     IX(i,j,k) = c0 + i + P*j + PP*k
// with P = roundup(Ni+1+4, 8) (i fastest, rows 64-byte aligned so that the
// first active cell i=1 sits on a 64-byte boundary), PP = P*(Nj+1).
// Cell (i,j,k), its low I/J/K faces and its low edges share the index, as
// in the reference's (isg:ieg[+1], ...) Fortran arrays.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GAM_TINY 1.0e-12
#define GAM_HUGE 1.0e+12
// sqrt(2.0) of the Boris "nuke" threshold (stress.F90:405): the intrinsic of a
// default-real argument returns default real, so the reference divides by the
// single-precision sqrt(2)
#define GAM_SQRT2_SP 1.41421353816986083984375

struct Geo {
    int P;          // i pitch
    long PP;        // k pitch
    long c0;        // index offset
    long N3;        // plane length (padded to 64 doubles)
    int Nip, Njp, Nkp, Ni, Nj, Nk;
    int is, ie, js, je, ks, ke;
    int isg, ieg, jsg, jeg, ksg, keg;
};

struct Box {
    int i0, i1, j0, j1, k0, k1;
};
// the active EMF edges of each edge direction
struct EdgeBoxes { Box e[3]; };

struct MParm {
    double gamma, CFL, Vd0, Ca, cHogH, cHogM, pdmb, dFloor, pFloor;
    double tFin, dt0, limDT0;
    double igm1;    // 1/(gamma-1)
    double iCa2;    // 1/Ca^2
    int doMHD, doBoris, doBG, doLFM, rmeth;
};

// Device-resident time/dt bookkeeping (Model_T%t/dt/ts, State%time, oState%time)
// Inflow series of the wind BC on the device (readWind, wind.F90:685-699)
struct WindDev {
    const double *tW, *Q, *B;  // n, (n,5), (n,3) column-major
    int n;
    double tMin, tMax, dtW;
    double Bx0, ByC, BzC;
    double xW, yW, zW;
};

struct Clock {
    double t;       // Model%t
    double dt;      // Model%dt
    double dtA;     // dt argument of AdvanceMHD
    double st;      // State%time
    double ost;     // oState%time
    double dtRet;   // last CalcDT result (standalone calc)
    unsigned long long dtBits;  // min-reduction accumulator (bits of a positive double)
    int ts;         // Model%ts
    int status;     // sticky CalcDT status: 0 ok, 1 bad, 2 disaster
    int lastStatus; // status of the last CalcDT
    int pad;
};

// Plane indices: State buffer
enum { S_GAS = 0, S_MF = 5, S_B = 8, S_E = 11, S_NPL = 14 };
// Half state buffer + primitive planes
enum { H_GAS = 0, H_MF = 5, H_B = 8, H_WH = 11, H_WN = 16, H_NPL = 21 };
// Metric planes
enum {
    M_XYZCC = 0, M_XFC = 3, M_DI = 12, M_DJ = 13, M_DK = 14, M_VOL = 15,
    M_TF = 16, M_TE = 43, M_TEB = 70, M_FACE = 82, M_EDGE = 85, M_XYZ = 88,
    M_NPL = 91
};
// Background planes
enum { B_FCB0 = 0, B_EDGB0 = 9, B_B0 = 15, B_BFLUX0 = 18, B_DPB0 = 21, B_NPL = 24 };

__host__ __device__ inline long IX(const Geo &g, int i, int j, int k)
{
    return g.c0 + (long)i + (long)g.P * j + g.PP * k;
}

template <int D> __host__ __device__ inline long STRIDE(const Geo &g)
{
    return D == 0 ? 1L : (D == 1 ? (long)g.P : g.PP);
}

// Load the double at byte offset `off` (32 bits: planes are far below
// 4 GiB) from a plane base that is uniform across the wavefront.  Written
// this way the compiler emits global loads with a scalar base and a 32-bit
// vector offset: one offset register serves every plane of a stencil
// point, with no 64-bit address arithmetic per load.
__device__ __forceinline__ double ldp(const double *__restrict__ plane, unsigned off)
{
    return *(const double *)((const char *)plane + off);
}
__device__ __forceinline__ void stp(double *__restrict__ plane, unsigned off, double v)
{
    *(double *)((char *)plane + off) = v;
}
// Point c + d*s of a plane whose stencil runs along axis AX (stride s):
// along i (AX = 0) plain indexing, which the compiler turns into wide loads
// of contiguous points; along j and k a 32-bit byte offset from the plane
// base (see ldp), the same for every plane at that point.
template <int AX>
__device__ __forceinline__ double ldx(const double *__restrict__ plane, long c, long s, int d)
{
    if (AX == 0) return plane[c + d];
    return ldp(plane, (unsigned)((c + d * s) * 8));
}

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

// Interpolation weights (recon.F90:22-32): integer numerators over a
// common denominator, rounded once each.
__device__ __forceinline__ double W8(int n)
{
    switch (n) {
    case 0: return -3.0 / 840.0;   case 1: return 29.0 / 840.0;
    case 2: return -139.0 / 840.0; case 3: return 533.0 / 840.0;
    case 4: return 533.0 / 840.0;  case 5: return -139.0 / 840.0;
    case 6: return 29.0 / 840.0;   default: return -3.0 / 840.0;
    }
}
__device__ __forceinline__ double W6(int n)
{
    switch (n) {
    case 0: return 0.0;          case 1: return 1.0 / 60.0;
    case 2: return -8.0 / 60.0;  case 3: return 37.0 / 60.0;
    case 4: return 37.0 / 60.0;  case 5: return -8.0 / 60.0;
    case 6: return 1.0 / 60.0;   default: return 0.0;
    }
}
__device__ __forceinline__ double UP7(int n)
{
    switch (n) {
    case 0: return -3.0 / 420.0;   case 1: return 25.0 / 420.0;
    case 2: return -101.0 / 420.0; case 3: return 319.0 / 420.0;
    case 4: return 214.0 / 420.0;  case 5: return -38.0 / 420.0;
    default: return 4.0 / 420.0;
    }
}

template <typename F> __device__ __forceinline__ double wsum8(F f)
{
    double s = 0.0;
#pragma unroll
    for (int n = 0; n < 8; n++) s += W8(n) * f(n);
    return s;
}
