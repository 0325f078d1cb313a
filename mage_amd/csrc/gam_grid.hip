// gam_grid.hip -- device pre-compute of the Grid_T metrics and the dipole
// background field (one-time setup, not part of the timed step).
//   k_face_metrics   face areas, centroids, triads   Corners2Grid init.F90:843-905
//   k_cell_metrics   volumes, centroids, di/dj/dk    init.F90:907-957
//   k_edge_metrics   edge lengths, edge triads       init.F90:959-1035
//   k_teb            edge-B transforms               ebGeom init.F90:1075-1137
//   k_b0_cells       dipole face/edge/volume data    AddB0 background.F90:54-138
//   k_b0_dp          background stress divergence    background.F90:141-153
// Quadrature: 12-point Gauss-Legendre (quadrature.F90:9-26).
#include "gam_common.h"
#include "gam_kernels.h"

__device__ __constant__ double cgA[12] = {
    0.1252334085114689154724, 0.3678314989981801937527,
    0.5873179542866174472967, 0.7699026741943046870369,
    0.9041172563704748566785, 0.9815606342467192506906,
   -0.1252334085114689154724,-0.3678314989981801937527,
   -0.5873179542866174472967,-0.7699026741943046870369,
   -0.9041172563704748566785,-0.9815606342467192506906 };
__device__ __constant__ double cgW[12] = {
    0.2491470458134027850006, 0.233492536538354808761,
    0.203167426723065921749,  0.160078328543346226335,
    0.1069393259953184309603, 0.0471753363865118271946,
    0.2491470458134027850006, 0.233492536538354808761,
    0.203167426723065921749,  0.160078328543346226335,
    0.1069393259953184309603, 0.0471753363865118271946 };

struct V3 { double x, y, z; };
__device__ __forceinline__ V3 v3(double a, double b, double c) { V3 r = {a, b, c}; return r; }
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 vcross(V3 a, V3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double vnorm(V3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ V3 vdiv(V3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }

__device__ __forceinline__ V3 corner(const Geo &g, const double *met, int i, int j, int k)
{
    const long c = IX(g, i, j, k);
    return v3(met[(M_XYZ + 0) * g.N3 + c], met[(M_XYZ + 1) * g.N3 + c], met[(M_XYZ + 2) * g.N3 + c]);
}

// Dipole (msphutils.F90:534-551), moment (0,0,M0)
__device__ __forceinline__ V3 dipole(double M0, V3 r)
{
    const double rad = vnorm(r);
    const double md = 0.0 * r.x + 0.0 * r.y + M0 * r.z;
    const double r5 = pow(rad, 5.0), r3 = pow(rad, 3.0);
    return v3(3 * md * r.x / r5 - 0.0 / r3, 3 * md * r.y / r5 - 0.0 / r3, 3 * md * r.z / r5 - M0 / r3);
}

// faceCoords (metric.F90:75-104)
__device__ __forceinline__ void face_corners(const Geo &g, const double *met, int d, int i, int j, int k,
                                             V3 &f0, V3 &f1, V3 &f2, V3 &f3)
{
    f0 = corner(g, met, i, j, k);
    if (d == 0) { f1 = corner(g, met, i, j + 1, k); f2 = corner(g, met, i, j, k + 1); f3 = corner(g, met, i, j + 1, k + 1); }
    else if (d == 1) { f1 = corner(g, met, i, j, k + 1); f2 = corner(g, met, i + 1, j, k); f3 = corner(g, met, i + 1, j, k + 1); }
    else { f1 = corner(g, met, i + 1, j, k); f2 = corner(g, met, i, j + 1, k); f3 = corner(g, met, i + 1, j + 1, k); }
}

// Bilinear face quadrature point (quadrature.F90:120-145)
struct FacePt { V3 x; V3 cr; double w; };
__device__ __forceinline__ FacePt face_point(V3 f0, V3 dx, V3 dy, V3 dz, int a, int b)
{
    // dx = (f1-f0, f2-f0, f3+f0-f2-f1) for the x component, etc. (faceDeltas)
    const double eta = 0.5 * (1 + cgA[a]), psi = 0.5 * (1 + cgA[b]);
    FacePt p;
    p.w = cgW[a] * cgW[b];
    p.x = v3(f0.x + dx.x * eta + dx.y * psi + dx.z * eta * psi,
             f0.y + dy.x * eta + dy.y * psi + dy.z * eta * psi,
             f0.z + dz.x * eta + dz.y * psi + dz.z * eta * psi);
    V3 ev = v3(dx.x + dx.z * psi, dy.x + dy.z * psi, dz.x + dz.z * psi);
    V3 pv = v3(dx.y + dx.z * eta, dy.y + dy.z * eta, dz.y + dz.z * eta);
    p.cr = vcross(ev, pv);
    return p;
}

__device__ __forceinline__ void face_deltas(V3 f0, V3 f1, V3 f2, V3 f3, V3 &dx, V3 &dy, V3 &dz)
{   // quadrature.F90:371-389
    dx = v3(f1.x - f0.x, f2.x - f0.x, f3.x + f0.x - f2.x - f1.x);
    dy = v3(f1.y - f0.y, f2.y - f0.y, f3.y + f0.y - f2.y - f1.y);
    dz = v3(f1.z - f0.z, f2.z - f0.z, f3.z + f0.z - f2.z - f1.z);
}

template <int D>
__global__ void k_face_metrics(Geo g, Box b, double *__restrict__ met)
{
    const int i = b.i0 + blockIdx.x * 64 + threadIdx.x;
    const int j = b.j0 + blockIdx.y;
    const int k = b.k0 + blockIdx.z;
    if (i > b.i1 || j > b.j1) return;
    const long N3 = g.N3, c = IX(g, i, j, k);
    V3 f0, f1, f2, f3, dx, dy, dz;
    face_corners(g, met, D, i, j, k, f0, f1, f2, f3);
    face_deltas(f0, f1, f2, f3, dx, dy, dz);
    double A = 0.0, Fx = 0.0, Fy = 0.0, Fz = 0.0;
    for (int a = 0; a < 12; a++)
        for (int bb = 0; bb < 12; bb++) {
            FacePt p = face_point(f0, dx, dy, dz, a, bb);
            const double ar = 0.25 * vnorm(p.cr) * p.w;
            A = A + ar * 1.0;
            Fx = Fx + ar * p.x.x; Fy = Fy + ar * p.x.y; Fz = Fz + ar * p.x.z;
        }
    met[(M_XFC + 0 + 3 * D) * N3 + c] = Fx / A;
    met[(M_XFC + 1 + 3 * D) * N3 + c] = Fy / A;
    met[(M_XFC + 2 + 3 * D) * N3 + c] = Fz / A;
    met[(M_FACE + D) * N3 + c] = A;
    // Face triad from corner differences (init.F90:871-898)
    V3 t1, t2;
    if (D == 0) {
        V3 a = corner(g, met, i, j + 1, k + 1), bq = corner(g, met, i, j, k + 1);
        V3 cq = corner(g, met, i, j + 1, k), e = corner(g, met, i, j, k);
        t1 = v3(0.5 * (a.x - bq.x) + 0.5 * (cq.x - e.x), 0.5 * (a.y - bq.y) + 0.5 * (cq.y - e.y), 0.5 * (a.z - bq.z) + 0.5 * (cq.z - e.z));
        t2 = v3(0.5 * (a.x - cq.x) + 0.5 * (bq.x - e.x), 0.5 * (a.y - cq.y) + 0.5 * (bq.y - e.y), 0.5 * (a.z - cq.z) + 0.5 * (bq.z - e.z));
    } else if (D == 1) {
        V3 a = corner(g, met, i + 1, j, k + 1), bq = corner(g, met, i + 1, j, k);
        V3 cq = corner(g, met, i, j, k + 1), e = corner(g, met, i, j, k);
        t1 = v3(0.5 * (a.x - bq.x) + 0.5 * (cq.x - e.x), 0.5 * (a.y - bq.y) + 0.5 * (cq.y - e.y), 0.5 * (a.z - bq.z) + 0.5 * (cq.z - e.z));
        t2 = v3(0.5 * (a.x - cq.x) + 0.5 * (bq.x - e.x), 0.5 * (a.y - cq.y) + 0.5 * (bq.y - e.y), 0.5 * (a.z - cq.z) + 0.5 * (bq.z - e.z));
    } else {
        V3 a = corner(g, met, i + 1, j + 1, k), bq = corner(g, met, i, j + 1, k);
        V3 cq = corner(g, met, i + 1, j, k), e = corner(g, met, i, j, k);
        t1 = v3(0.5 * (a.x - bq.x) + 0.5 * (cq.x - e.x), 0.5 * (a.y - bq.y) + 0.5 * (cq.y - e.y), 0.5 * (a.z - bq.z) + 0.5 * (cq.z - e.z));
        t2 = v3(0.5 * (a.x - cq.x) + 0.5 * (bq.x - e.x), 0.5 * (a.y - cq.y) + 0.5 * (bq.y - e.y), 0.5 * (a.z - cq.z) + 0.5 * (bq.z - e.z));
    }
    V3 cr = vcross(t1, t2);
    V3 eN = vdiv(cr, vnorm(cr));
    V3 eT2 = vdiv(t2, vnorm(t2));
    V3 c2 = vcross(eT2, eN);
    V3 eT1 = vdiv(c2, vnorm(c2));
    double *Tf = met + (long)(M_TF + 9 * D) * N3 + c;
    Tf[0] = eN.x; Tf[N3] = eN.y; Tf[2 * N3] = eN.z;
    Tf[3 * N3] = eT1.x; Tf[4 * N3] = eT1.y; Tf[5 * N3] = eT1.z;
    Tf[6 * N3] = eT2.x; Tf[7 * N3] = eT2.y; Tf[8 * N3] = eT2.z;
}

// GaussianVolumeIntegral (quadrature.F90:42-96) over cell (i,j,k).
// mode 0: integrand (1,x,y,z); mode 1: dipole B.  out[0..3]
template <int MODE>
__device__ void vol_quad(const Geo &g, const double *met, int i, int j, int k, double M0, double *out)
{
    V3 X[8];
    X[0] = corner(g, met, i, j, k);         X[1] = corner(g, met, i + 1, j, k);
    X[2] = corner(g, met, i + 1, j + 1, k); X[3] = corner(g, met, i, j + 1, k);
    X[4] = corner(g, met, i, j, k + 1);     X[5] = corner(g, met, i + 1, j, k + 1);
    X[6] = corner(g, met, i + 1, j + 1, k + 1); X[7] = corner(g, met, i, j + 1, k + 1);
    const int iS[8] = {-1, +1, +1, -1, -1, +1, +1, -1};
    const int jS[8] = {-1, -1, +1, +1, -1, -1, +1, +1};
    const int kS[8] = {-1, -1, -1, -1, +1, +1, +1, +1};
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int a = 0; a < 12; a++)
        for (int b = 0; b < 12; b++)
            for (int cc = 0; cc < 12; cc++) {
                const double wijk = cgW[a] * cgW[b] * cgW[cc];
                double xp[3], J[3][3];
                for (int x = 0; x < 3; x++) {
                    double s = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        const double is = (iS[q] < 0) ? 1 - cgA[a] : 1 + cgA[a];
                        const double js = (jS[q] < 0) ? 1 - cgA[b] : 1 + cgA[b];
                        const double ks = (kS[q] < 0) ? 1 - cgA[cc] : 1 + cgA[cc];
                        const double xq = (x == 0) ? X[q].x : (x == 1) ? X[q].y : X[q].z;
                        s = s + xq * (is * js * ks);
                        s1 = s1 + xq * (iS[q] * js * ks);
                        s2 = s2 + xq * (is * jS[q] * ks);
                        s3 = s3 + xq * (is * js * kS[q]);
                    }
                    xp[x] = 0.125 * s;
                    J[x][0] = 0.125 * s1; J[x][1] = 0.125 * s2; J[x][2] = 0.125 * s3;
                }
                double dJ = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
                          - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
                          + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                dJ = fabs(dJ);
                double Wp[4];
                if (MODE == 0) { Wp[0] = 1.0; Wp[1] = xp[0]; Wp[2] = xp[1]; Wp[3] = xp[2]; }
                else { V3 B = dipole(M0, v3(xp[0], xp[1], xp[2])); Wp[0] = B.x; Wp[1] = B.y; Wp[2] = B.z; Wp[3] = 0.0; }
                for (int v = 0; v < 4; v++) acc[v] = acc[v] + dJ * wijk * Wp[v];
            }
    for (int v = 0; v < 4; v++) out[v] = acc[v];
}

__global__ void k_cell_metrics(Geo g, Box b, double *__restrict__ met)
{
    const int i = b.i0 + blockIdx.x * 64 + threadIdx.x;
    const int j = b.j0 + blockIdx.y;
    const int k = b.k0 + blockIdx.z;
    if (i > b.i1 || j > b.j1) return;
    const long N3 = g.N3, c = IX(g, i, j, k);
    double w[4];
    vol_quad<0>(g, met, i, j, k, 0.0, w);
    met[M_VOL * N3 + c] = w[0];
    for (int x = 0; x < 3; x++) met[(M_XYZCC + x) * N3 + c] = w[1 + x] / w[0];
    double fa;
    fa = fmax(met[(M_FACE + 0) * N3 + c + 1], met[(M_FACE + 0) * N3 + c]);
    met[M_DI * N3 + c] = w[0] / fa;
    fa = fmax(met[(M_FACE + 1) * N3 + c + g.P], met[(M_FACE + 1) * N3 + c]);
    met[M_DJ * N3 + c] = w[0] / fa;
    fa = fmax(met[(M_FACE + 2) * N3 + c + g.PP], met[(M_FACE + 2) * N3 + c]);
    met[M_DK * N3 + c] = w[0] / fa;
}

template <int D>
__global__ void k_edge_metrics(Geo g, Box b, double *__restrict__ met)
{
    const int i = b.i0 + blockIdx.x * 64 + threadIdx.x;
    const int j = b.j0 + blockIdx.y;
    const int k = b.k0 + blockIdx.z;
    if (i > b.i1 || j > b.j1) return;
    const long N3 = g.N3, c = IX(g, i, j, k);
    V3 p0, p1, a0, a1, b0, b1;
    if (D == 0) {
        p1 = corner(g, met, i + 1, j, k); p0 = corner(g, met, i, j, k);
        a1 = corner(g, met, i, j, k + 1); a0 = corner(g, met, i, j, k - 1);
        b1 = corner(g, met, i + 1, j, k + 1); b0 = corner(g, met, i + 1, j, k - 1);
    } else if (D == 1) {
        p1 = corner(g, met, i, j + 1, k); p0 = corner(g, met, i, j, k);
        a1 = corner(g, met, i + 1, j, k); a0 = corner(g, met, i - 1, j, k);
        b1 = corner(g, met, i + 1, j + 1, k); b0 = corner(g, met, i - 1, j + 1, k);
    } else {
        p1 = corner(g, met, i, j, k + 1); p0 = corner(g, met, i, j, k);
        a1 = corner(g, met, i, j + 1, k); a0 = corner(g, met, i, j - 1, k);
        b1 = corner(g, met, i, j + 1, k + 1); b0 = corner(g, met, i, j - 1, k + 1);
    }
    V3 dE = vsub(p1, p0);
    V3 eA = v3(0.5 * (a1.x - a0.x + b1.x - b0.x), 0.5 * (a1.y - a0.y + b1.y - b0.y), 0.5 * (a1.z - a0.z + b1.z - b0.z));
    const double nE = vnorm(dE);
    V3 eN = vdiv(dE, nE);
    V3 c1 = vcross(eA, eN);
    V3 eT1 = vdiv(c1, vnorm(c1));
    V3 c2 = vcross(eN, eT1);
    V3 eT2 = vdiv(c2, vnorm(c2));
    met[(M_EDGE + D) * N3 + c] = nE;
    double *Te = met + (long)(M_TE + 9 * D) * N3 + c;
    Te[0] = eN.x; Te[N3] = eN.y; Te[2 * N3] = eN.z;
    Te[3 * N3] = eT1.x; Te[4 * N3] = eT1.y; Te[5 * N3] = eT1.z;
    Te[6 * N3] = eT2.x; Te[7 * N3] = eT2.y; Te[8 * N3] = eT2.z;
}

// ebGeom (init.F90:1075-1137): FD faces stacked along SD, mapped into the
// (TAN1,TAN2) system of edge direction DN; writes components C0, C0+1.
template <int DN, int FD, int SD, int C0>
__global__ void k_teb(Geo g, Box b, double *__restrict__ met)
{
    const int i = b.i0 + blockIdx.x * 64 + threadIdx.x;
    const int j = b.j0 + blockIdx.y;
    const int k = b.k0 + blockIdx.z;
    if (i > b.i1 || j > b.j1) return;
    const long N3 = g.N3, c = IX(g, i, j, k);
    const long s = STRIDE<SD>(g);
    double fA[8], Nx[8], Ny[8], Nz[8];
    for (int n = 0; n < 8; n++) {
        const long p = c + (n - 4) * s;
        fA[n] = met[(M_FACE + FD) * N3 + p];
        Nx[n] = met[(M_TF + 0 + 9 * FD) * N3 + p];
        Ny[n] = met[(M_TF + 1 + 9 * FD) * N3 + p];
        Nz[n] = met[(M_TF + 2 + 9 * FD) * N3 + p];
    }
    double fAi = 0.0, ex = 0.0, ey = 0.0, ez = 0.0;
    for (int n = 0; n < 8; n++) {
        fAi = fAi + W6(n) * fA[n];
        ex = ex + (Nx[n] * fA[n]) * W6(n);
        ey = ey + (Ny[n] * fA[n]) * W6(n);
        ez = ez + (Nz[n] * fA[n]) * W6(n);
    }
    ex = ex / fAi; ey = ey / fAi; ez = ez / fAi;
    const double *Te = met + (long)(M_TE + 9 * DN) * N3 + c;
    const double dxT = ex * Te[3 * N3] + ey * Te[4 * N3] + ez * Te[5 * N3];
    const double dyT = ex * Te[6 * N3] + ey * Te[7 * N3] + ez * Te[8 * N3];
    double dS = sqrt(dxT * dxT + dyT * dyT);
    if (dS <= GAM_TINY) dS = 1.0;
    met[(M_TEB + C0 + 4 * DN) * N3 + c] = dxT / dS;
    met[(M_TEB + C0 + 1 + 4 * DN) * N3 + c] = dyT / dS;
}

__global__ void k_fill(long n, double v, double *__restrict__ p)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = v;
}

// AddB0 (background.F90:54-138) per cell for the dipole
__global__ void k_b0_cells(Geo g, Box b, double M0, const double *__restrict__ met,
                           double *__restrict__ bg, double *__restrict__ fs)
{
    const int i = b.i0 + blockIdx.x * 64 + threadIdx.x;
    const int j = b.j0 + blockIdx.y;
    const int k = b.k0 + blockIdx.z;
    if (i > b.i1 || j > b.j1) return;
    const long N3 = g.N3, c = IX(g, i, j, k);
    for (int d = 0; d < 3; d++) {
        V3 f0, f1, f2, f3, dx, dy, dz;
        face_corners(g, met, d, i, j, k, f0, f1, f2, f3);
        face_deltas(f0, f1, f2, f3, dx, dy, dz);
        double flx = 0.0, Fx = 0.0, Fy = 0.0, Fz = 0.0, Sx = 0.0, Sy = 0.0, Sz = 0.0;
        for (int a = 0; a < 12; a++)
            for (int bb = 0; bb < 12; bb++) {
                FacePt p = face_point(f0, dx, dy, dz, a, bb);
                V3 B = dipole(M0, p.x);
                const double ar = 0.25 * vnorm(p.cr) * p.w;
                Fx = Fx + ar * B.x; Fy = Fy + ar * B.y; Fz = Fz + ar * B.z;
                const double qw = 0.25 * p.w;
                V3 N = v3(qw * p.cr.x, qw * p.cr.y, qw * p.cr.z);
                flx = flx + vdot(N, B);
                const double bb2 = vdot(B, B), bdS = vdot(B, N);
                Sx = Sx + 0.5 * bb2 * N.x - B.x * bdS;
                Sy = Sy + 0.5 * bb2 * N.y - B.y * bdS;
                Sz = Sz + 0.5 * bb2 * N.z - B.z * bdS;
            }
        const double fa = met[(M_FACE + d) * N3 + c];
        bg[(B_BFLUX0 + d) * N3 + c] = flx;
        bg[(B_FCB0 + 0 + 3 * d) * N3 + c] = Fx / fa;
        bg[(B_FCB0 + 1 + 3 * d) * N3 + c] = Fy / fa;
        bg[(B_FCB0 + 2 + 3 * d) * N3 + c] = Fz / fa;
        fs[(d + 0) * N3 + c] = Sx;
        fs[(d + 3) * N3 + c] = Sy;
        fs[(d + 6) * N3 + c] = Sz;
    }
    for (int d = 0; d < 3; d++) {
        V3 e1 = corner(g, met, i, j, k);
        V3 e2 = corner(g, met, i + (d == 0), j + (d == 1), k + (d == 2));
        V3 r2 = v3(0.5 * (e2.x - e1.x), 0.5 * (e2.y - e1.y), 0.5 * (e2.z - e1.z));
        V3 xb = v3(0.5 * (e2.x + e1.x), 0.5 * (e2.y + e1.y), 0.5 * (e2.z + e1.z));
        double ex = 0.0, ey = 0.0, ez = 0.0;
        for (int q = 0; q < 12; q++) {
            V3 xp = v3(xb.x + cgA[q] * r2.x, xb.y + cgA[q] * r2.y, xb.z + cgA[q] * r2.z);
            V3 A = dipole(M0, xp);
            const double hw = 0.5 * cgW[q];
            ex = ex + hw * A.x; ey = ey + hw * A.y; ez = ez + hw * A.z;
        }
        const double *Te = met + (long)(M_TE + 9 * d) * N3 + c;
        bg[(B_EDGB0 + 0 + 2 * d) * N3 + c] = ex * Te[3 * N3] + ey * Te[4 * N3] + ez * Te[5 * N3];
        bg[(B_EDGB0 + 1 + 2 * d) * N3 + c] = ex * Te[6 * N3] + ey * Te[7 * N3] + ez * Te[8 * N3];
    }
    double w[4];
    vol_quad<1>(g, met, i, j, k, M0, w);
    const double vol = met[M_VOL * N3 + c];
    for (int x = 0; x < 3; x++) bg[(B_B0 + x) * N3 + c] = w[x] / vol;
}

__global__ void k_b0_dp(Geo g, Box b, const double *__restrict__ met, double *__restrict__ bg,
                        const double *__restrict__ fs)
{
    const int i = b.i0 + blockIdx.x * 64 + threadIdx.x;
    const int j = b.j0 + blockIdx.y;
    const int k = b.k0 + blockIdx.z;
    if (i > b.i1 || j > b.j1) return;
    const long N3 = g.N3, c = IX(g, i, j, k);
    for (int x = 0; x < 3; x++) {
        const double *f = fs + (long)(3 * x) * N3;
        bg[(B_DPB0 + x) * N3 + c] = -(f[c + 1] - f[c] + f[N3 + c + g.P] - f[N3 + c]
                                      + f[2 * N3 + c + g.PP] - f[2 * N3 + c]) / met[M_VOL * N3 + c];
    }
}

static inline dim3 gr3(const Box &b)
{
    return dim3((unsigned)((b.i1 - b.i0 + 64) / 64), (unsigned)(b.j1 - b.j0 + 1), (unsigned)(b.k1 - b.k0 + 1));
}
static inline Box mkbox(int i0, int i1, int j0, int j1, int k0, int k1)
{
    Box b = {i0, i1, j0, j1, k0, k1};
    return b;
}

static void fill(hipStream_t s, double *p, long n, double v)
{
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, v, p);
}

void launch_grid_from_corners(hipStream_t s, const Geo &g, double *met)
{
    const long N3 = g.N3;
    // Defaults as in Corners2Grid: face=1, xfc=0, Tf=0, edge=1, Te=0, Teb=1
    fill(s, met + (long)M_FACE * N3, 3 * N3, 1.0);
    fill(s, met + (long)M_EDGE * N3, 3 * N3, 1.0);
    fill(s, met + (long)M_XFC * N3, 9 * N3, 0.0);
    fill(s, met + (long)M_TF * N3, 27 * N3, 0.0);
    fill(s, met + (long)M_TE * N3, 27 * N3, 0.0);
    fill(s, met + (long)M_TEB * N3, 12 * N3, 1.0);
    Box b;
    b = mkbox(g.isg, g.ieg + 1, g.jsg, g.jeg, g.ksg, g.keg);
    hipLaunchKernelGGL(k_face_metrics<0>, gr3(b), dim3(64), 0, s, g, b, met);
    b = mkbox(g.isg, g.ieg, g.jsg, g.jeg + 1, g.ksg, g.keg);
    hipLaunchKernelGGL(k_face_metrics<1>, gr3(b), dim3(64), 0, s, g, b, met);
    b = mkbox(g.isg, g.ieg, g.jsg, g.jeg, g.ksg, g.keg + 1);
    hipLaunchKernelGGL(k_face_metrics<2>, gr3(b), dim3(64), 0, s, g, b, met);
    b = mkbox(g.isg, g.ieg, g.jsg, g.jeg, g.ksg, g.keg);
    hipLaunchKernelGGL(k_cell_metrics, gr3(b), dim3(64), 0, s, g, b, met);
    b = mkbox(g.isg, g.ieg, g.jsg, g.jeg + 1, g.ksg + 1, g.keg);
    hipLaunchKernelGGL(k_edge_metrics<0>, gr3(b), dim3(64), 0, s, g, b, met);
    b = mkbox(g.isg + 1, g.ieg, g.jsg, g.jeg, g.ksg, g.keg + 1);
    hipLaunchKernelGGL(k_edge_metrics<1>, gr3(b), dim3(64), 0, s, g, b, met);
    b = mkbox(g.isg, g.ieg + 1, g.jsg + 1, g.jeg, g.ksg, g.keg);
    hipLaunchKernelGGL(k_edge_metrics<2>, gr3(b), dim3(64), 0, s, g, b, met);
    b = mkbox(g.is, g.ie + 1, g.js, g.je + 1, g.ks, g.ke + 1);
    // dNorm = I: (T1,T2) = (J,K); J: (K,I); K: (I,J)
    hipLaunchKernelGGL((k_teb<0, 1, 2, 0>), gr3(b), dim3(64), 0, s, g, b, met);
    hipLaunchKernelGGL((k_teb<0, 2, 1, 2>), gr3(b), dim3(64), 0, s, g, b, met);
    hipLaunchKernelGGL((k_teb<1, 2, 0, 0>), gr3(b), dim3(64), 0, s, g, b, met);
    hipLaunchKernelGGL((k_teb<1, 0, 2, 2>), gr3(b), dim3(64), 0, s, g, b, met);
    hipLaunchKernelGGL((k_teb<2, 0, 1, 0>), gr3(b), dim3(64), 0, s, g, b, met);
    hipLaunchKernelGGL((k_teb<2, 1, 0, 2>), gr3(b), dim3(64), 0, s, g, b, met);
}

void launch_dipole_b0(hipStream_t s, const Geo &g, double M0, const double *met, double *bg,
                      double *scratch9)
{
    const long N3 = g.N3;
    fill(s, bg, (long)B_NPL * N3, 0.0);
    fill(s, scratch9, 9 * N3, 0.0);
    Box b = mkbox(g.isg, g.ieg, g.jsg, g.jeg, g.ksg, g.keg);
    hipLaunchKernelGGL(k_b0_cells, gr3(b), dim3(64), 0, s, g, b, M0, met, bg, scratch9);
    b = mkbox(g.is, g.ie, g.js, g.je, g.ks, g.ke);
    hipLaunchKernelGGL(k_b0_dp, gr3(b), dim3(64), 0, s, g, b, met, bg, scratch9);
}
