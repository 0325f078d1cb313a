// gam_adj.hip -- gfx950 kernels of the discrete adjoint of the
// hydrodynamic GAMERA step (doMHD = 0): the transposes of the boundary
// copies, of the Reynolds update (Flux2Deltas + applyUpdate,
// mhdgroup.F90:79-121, stress.F90:152-194), of the interface fluxes
// (Fluxes, stress.F90:204-429: reconstruction, PDM, GasKinFlux) and of the
// predictor (mhdgroup.F90:444-561).
//
// Adjoint planes hold the derivative of the objective with respect to the
// same-named forward plane.  The flux adjoint scatters into the stencil
// cells with FP64 atomic adds (hardware global_atomic_add_f64); every other
// kernel gathers.  dt is the forward run's value at each step (the CFL
// minimum is held fixed, as the finite-difference checks also do).
#include "gam_adj.h"
#include "gam_kernels.h"

#define ABX 64
#define ABY 4

static inline dim3 agrid(const Box &b)
{
    return dim3((unsigned)((b.i1 - b.i0 + ABX) / ABX), (unsigned)((b.j1 - b.j0 + ABY) / ABY),
                (unsigned)(b.k1 - b.k0 + 1));
}

#define POINT_IN(b)                                          \
    const int i = (b).i0 + (int)(blockIdx.x * ABX + threadIdx.x); \
    const int j = (b).j0 + (int)(blockIdx.y * ABY + threadIdx.y); \
    const int k = (b).k0 + (int)blockIdx.z;                  \
    if (i > (b).i1 || j > (b).j1) return;

__device__ __forceinline__ bool inside(const Box &b, int i, int j, int k)
{
    return i >= b.i0 && i <= b.i1 && j >= b.j0 && j <= b.j1 && k >= b.k0 && k <= b.k1;
}

// AB2 weight of the predictor from a step's clock (k_predict)
__device__ __forceinline__ double ab2_weight(const Clock *clk)
{
    const double st = clk->st, dtM = clk->dt;
    double ost = clk->ost;
    if (fabs(ost - st) < GAM_TINY) ost = st - dtM;
    return (0.5 * clk->dtA) / (st - ost);
}

// ------------------------------------------------------------ objective
// J = sum_c w_c P_c over cells (P from CellC2P); seeds the adjoint of the
// final State's gas and writes one partial sum per block.
__global__ __launch_bounds__(256) void k_adj_objective(Geo g, MParm m, Box cb,
                                                       const double *__restrict__ S,
                                                       const double *__restrict__ wgt,
                                                       double *__restrict__ aS,
                                                       double *__restrict__ part)
{
    __shared__ double red[ABX * ABY];
    const int tid = threadIdx.x + ABX * threadIdx.y;
    const int i = cb.i0 + (int)(blockIdx.x * ABX + threadIdx.x);
    const int j = cb.j0 + (int)(blockIdx.y * ABY + threadIdx.y);
    const int k = cb.k0 + (int)blockIdx.z;
    double jv = 0.0;
    if (i <= cb.i1 && j <= cb.j1) {
        const long N3 = g.N3;
        const long c = IX(g, i, j, k);
        double U[5], W[5], aW[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, aU[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int v = 0; v < 5; v++) U[v] = S[(S_GAS + v) * N3 + c];
        c2p_cell(m, U, W);
        const double w = wgt[c];
        jv = w * W[PRS];
        aW[PRS] = w;
        c2p_cell_adj(m, U, aW, aU);
        for (int v = 0; v < 5; v++) aS[v * N3 + c] = aU[v];
    }
    red[tid] = jv;
    __syncthreads();
    for (int h = ABX * ABY / 2; h > 0; h >>= 1) {
        if (tid < h) red[tid] += red[tid + h];
        __syncthreads();
    }
    if (tid == 0) part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = red[0];
}

// ------------------------------------------------------------ boundaries
// Transpose of one side's gas copies (k_bc_periodic / k_bc_copy): every
// ghost's adjoint moves onto the cell it was copied from.  One thread per
// transverse point handles all four layers, so a source shared by several
// layers (zero gradient) is updated by one thread.
template <int SIDE, int KIND>
__global__ void k_bc_adj(Geo g, Box t, double *__restrict__ aS)
{
    const int a = t.i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int b = t.j0 + (int)blockIdx.y;
    if (a > t.i1 || b > t.j1) return;
    const long N3 = g.N3;
    for (int n = 1; n <= 4; n++) {
        int gi, gj, gk, si, sj, sk;
        if (SIDE <= 1) {
            gj = sj = a; gk = sk = b;
            gi = SIDE == 0 ? g.is - n : g.ie + n;
            si = KIND == 0 ? (SIDE == 0 ? g.ie - n + 1 : g.is + n - 1) : (SIDE == 0 ? g.is : g.ie);
        } else if (SIDE <= 3) {
            gi = si = a; gk = sk = b;
            gj = SIDE == 2 ? g.js - n : g.je + n;
            sj = KIND == 0 ? (SIDE == 2 ? g.je - n + 1 : g.js + n - 1) : (SIDE == 2 ? g.js : g.je);
        } else {
            gi = si = a; gj = sj = b;
            gk = SIDE == 4 ? g.ks - n : g.ke + n;
            sk = KIND == 0 ? (SIDE == 4 ? g.ke - n + 1 : g.ks + n - 1) : (SIDE == 4 ? g.ks : g.ke);
        }
        if (gi > g.ieg || gj > g.jeg || gk > g.keg) continue;   // face-only points
        const long d = IX(g, gi, gj, gk), s = IX(g, si, sj, sk);
        for (int v = 0; v < 5; v++) {
            aS[v * N3 + s] += aS[v * N3 + d];
            aS[v * N3 + d] = 0.0;
        }
    }
}

// ------------------------------------------------------------ update
// Transpose of Sn = So + dtM * dH (active cells), Sn = So (other cells):
// aSo += aSn everywhere; the adjoint of the active faces' fluxes is
// dtM (aSn/V of the cell above the face - aSn/V of the cell below).
__global__ __launch_bounds__(256) void k_update_adj(Geo g, Box fb, Box ab, const Clock *clk,
                                                    const double *__restrict__ aSn,
                                                    double *__restrict__ aSo,
                                                    double *__restrict__ ag,
                                                    const double *__restrict__ met)
{
    POINT_IN(fb);
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    if (i <= g.ieg && j <= g.jeg && k <= g.keg)
        for (int v = 0; v < 5; v++) aSo[v * N3 + c] += aSn[v * N3 + c];
    const double dtM = clk->dt;
    const long st[3] = {1, g.P, g.PP};
    for (int d = 0; d < 3; d++) {
        // active faces of direction d (the boxes the flux kernels cover)
        const bool face = i >= g.is && j >= g.js && k >= g.ks && i <= g.ie + (d == 0) &&
                          j <= g.je + (d == 1) && k <= g.ke + (d == 2);
        if (!face) continue;
        const long cl = c - st[d];
        const bool up = inside(ab, i, j, k);
        const bool lo = inside(ab, i - (d == 0), j - (d == 1), k - (d == 2));
        const double wu = up ? dtM / met[M_VOL * N3 + c] : 0.0;
        const double wl = lo ? dtM / met[M_VOL * N3 + cl] : 0.0;
        for (int v = 0; v < 5; v++)
            ag[(v + 5 * d) * N3 + c] = wu * aSn[v * N3 + c] - wl * aSn[v * N3 + cl];
    }
}

// ------------------------------------------------------------ fluxes
// Transpose of one direction's gas fluxes: recompute the face's L/R
// states from the half-state planes, then the adjoints of GasKinFlux and
// of the reconstruction, scattered into the stencil cells' adjoints of
// H_WH (volume-weighted half primitives) and H_WN (limiter primitives).
template <int D, int RM>
__global__ __launch_bounds__(256) void k_flux_adj(Geo g, MParm m, Box fb,
                                                  const double *__restrict__ H,
                                                  const double *__restrict__ met,
                                                  const double *__restrict__ ag,
                                                  double *__restrict__ aWH,
                                                  double *__restrict__ aWN)
{
    POINT_IN(fb);
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    const long s = STRIDE<D>(g);
    double vol[8];
    for (int n = 0; n < 8; n++) vol[n] = met[M_VOL * N3 + c + (n - 4) * s];
    const double rVi = 1.0 / wsum8([&](int n) { return vol[n]; });
    double qv[5][8], nq[5][4], WL[5], WR[5];
    for (int v = 0; v < 5; v++) {
        for (int n = 0; n < 8; n++) qv[v][n] = H[(H_WH + v) * N3 + c + (n - 4) * s];
        for (int n = 0; n < 4; n++) nq[v][n] = H[(H_WN + v) * N3 + c + (n - 2) * s];
        recon_lr_w<RM>(m.pdmb, rVi, nq[v], qv[v], WL[v], WR[v]);
    }
    double T[9];
    for (int q = 0; q < 9; q++) T[q] = met[(M_TF + q + 9 * D) * N3 + c];
    const double faB = met[(M_FACE + D) * N3 + c];
    double aF[5], aL[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, aR[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int v = 0; v < 5; v++) aF[v] = faB * ag[(v + 5 * D) * N3 + c];
    gas_kin_adj(m.igm1, -1.0, WL, T, aF, aL);
    gas_kin_adj(m.igm1, +1.0, WR, T, aF, aR);
    for (int v = 0; v < 5; v++) {
        double aqv[8] = {0, 0, 0, 0, 0, 0, 0, 0}, anq[4] = {0, 0, 0, 0};
        recon_lr_w_adj<RM>(m.pdmb, rVi, nq[v], qv[v], aL[v], aR[v], anq, aqv);
        for (int n = 0; n < 8; n++) unsafeAtomicAdd(aWH + v * N3 + c + (n - 4) * s, aqv[n]);
        for (int n = 0; n < 4; n++) unsafeAtomicAdd(aWN + v * N3 + c + (n - 2) * s, anq[n]);
    }
}

// ------------------------------------------------------------ predictor
// Transpose of the predictor's gas outputs at every cell: adds the adjoint
// of State(n) gas into aS, writes that of oState gas into aO.
__global__ __launch_bounds__(256) void k_predict_adj(Geo g, MParm m, Box cb, const Clock *clk,
                                                     const double *__restrict__ S,
                                                     const double *__restrict__ O,
                                                     const double *__restrict__ met,
                                                     const double *__restrict__ aWH,
                                                     const double *__restrict__ aWN,
                                                     double *__restrict__ aS,
                                                     double *__restrict__ aO)
{
    POINT_IN(cb);
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    const double w = ab2_weight(clk);
    double Q[5], oQ[5], aHG[5] = {0, 0, 0, 0, 0}, awh[5], awn[5], aQ[5] = {0, 0, 0, 0, 0},
           aoQ[5] = {0, 0, 0, 0, 0};
    for (int v = 0; v < 5; v++) {
        Q[v] = S[(S_GAS + v) * N3 + c];
        oQ[v] = O[(S_GAS + v) * N3 + c];
        awh[v] = aWH[v * N3 + c];
        awn[v] = aWN[v * N3 + c];
    }
    predict_cell_adj(m, Q, oQ, w, met[M_VOL * N3 + c], aHG, awh, awn, aQ, aoQ);
    for (int v = 0; v < 5; v++) {
        aS[v * N3 + c] += aQ[v];
        aO[v * N3 + c] = aoQ[v];
    }
}

// ------------------------------------------------------------ replay
// Hold dt at a recorded value (frozen-dt replays of the forward run).
__global__ void k_set_dt(Clock *clk, const Clock *rec)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Clock c = *clk;
    c.dt = rec->dt;
    c.dtA = rec->dtA;
    *clk = c;
}

// ------------------------------------------------------------ launchers
void launch_adj_objective(hipStream_t s, const Geo &g, const MParm &m, const double *S,
                          const double *wgt, double *aS, double *part, int *nparts)
{
    const Box cb = {g.isg, g.ieg, g.jsg, g.jeg, g.ksg, g.keg};
    const dim3 gr = agrid(cb);
    *nparts = (int)(gr.x * gr.y * gr.z);
    hipLaunchKernelGGL(k_adj_objective, gr, dim3(ABX, ABY), 0, s, g, m, cb, S, wgt, aS, part);
}

template <int SIDE>
static void bc_adj_side(hipStream_t s, int kind, const Geo &g, const Box &t, double *aS)
{
    const dim3 gr((unsigned)((t.i1 - t.i0 + 64) / 64), (unsigned)(t.j1 - t.j0 + 1));
    if (kind == 0) hipLaunchKernelGGL((k_bc_adj<SIDE, 0>), gr, dim3(64), 0, s, g, t, aS);
    else hipLaunchKernelGGL((k_bc_adj<SIDE, 1>), gr, dim3(64), 0, s, g, t, aS);
}

void launch_bc_adj(hipStream_t s, int side, int kind, const Geo &g, double *aS)
{
    // transverse ranges of the forward sides (launch_bc), cells only
    Box t;
    if (side <= 1) { t.i0 = g.jsg; t.i1 = g.jeg; t.j0 = g.ksg; t.j1 = g.keg; }
    else if (side <= 3) { t.i0 = g.isg; t.i1 = g.ieg; t.j0 = g.ksg; t.j1 = g.keg; }
    else { t.i0 = g.isg; t.i1 = g.ieg; t.j0 = g.jsg; t.j1 = g.jeg; }
    t.k0 = t.k1 = 0;
    switch (side) {
    case 0: bc_adj_side<0>(s, kind, g, t, aS); break;
    case 1: bc_adj_side<1>(s, kind, g, t, aS); break;
    case 2: bc_adj_side<2>(s, kind, g, t, aS); break;
    case 3: bc_adj_side<3>(s, kind, g, t, aS); break;
    case 4: bc_adj_side<4>(s, kind, g, t, aS); break;
    default: bc_adj_side<5>(s, kind, g, t, aS); break;
    }
}

void launch_update_adj(hipStream_t s, const Geo &g, const Box &ab, const Clock *clk, const double *aSn,
                       double *aSo, double *ag, const double *met)
{
    const Box fb = {g.isg, g.ieg + 1, g.jsg, g.jeg + 1, g.ksg, g.keg + 1};
    hipLaunchKernelGGL(k_update_adj, agrid(fb), dim3(ABX, ABY), 0, s, g, fb, ab, clk, aSn, aSo, ag, met);
}

template <int D>
static void flux_adj_d(hipStream_t s, const Geo &g, const MParm &m, const Box &fb, const double *H,
                       const double *met, const double *ag, double *aWH, double *aWN)
{
    if (m.rmeth) hipLaunchKernelGGL((k_flux_adj<D, 1>), agrid(fb), dim3(ABX, ABY), 0, s, g, m, fb, H, met, ag, aWH, aWN);
    else hipLaunchKernelGGL((k_flux_adj<D, 0>), agrid(fb), dim3(ABX, ABY), 0, s, g, m, fb, H, met, ag, aWH, aWN);
}

void launch_flux_adj(hipStream_t s, int d, const Geo &g, const MParm &m, const Box &fb, const double *H,
                     const double *met, const double *ag, double *aWH, double *aWN)
{
    if (d == 0) flux_adj_d<0>(s, g, m, fb, H, met, ag, aWH, aWN);
    else if (d == 1) flux_adj_d<1>(s, g, m, fb, H, met, ag, aWH, aWN);
    else flux_adj_d<2>(s, g, m, fb, H, met, ag, aWH, aWN);
}

void launch_predict_adj(hipStream_t s, const Geo &g, const MParm &m, const Clock *clk, const double *S,
                        const double *O, const double *met, const double *aWH, const double *aWN,
                        double *aS, double *aO)
{
    const Box cb = {g.isg, g.ieg, g.jsg, g.jeg, g.ksg, g.keg};
    hipLaunchKernelGGL(k_predict_adj, agrid(cb), dim3(ABX, ABY), 0, s, g, m, cb, clk, S, O, met, aWH, aWN,
                       aS, aO);
}

void launch_set_dt(hipStream_t s, Clock *clk, const Clock *rec)
{
    hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(64), 0, s, clk, rec);
}
