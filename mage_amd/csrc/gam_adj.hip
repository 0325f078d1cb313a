// gam_adj.hip -- gfx950 kernels of the discrete adjoint of the GAMERA
// step: the transposes of the boundary copies and of the ghost bFlux2Fld,
// of the update (Flux2Deltas, Reynolds, E2Flux/CT, cell B, CellMaxwell;
// mhdgroup.F90:79-329, stress.F90:152-194, fields.F90:18-51), of the
// interface fluxes (Fluxes, stress.F90:204-599: reconstruction, PDM,
// GasKinFlux, MagKinFlux, HOGS), of the edge EMF (CalcElecField,
// fields.F90:70-517), of Mom2Face and of the predictor
// (mhdgroup.F90:444-561).
//
// Adjoint planes hold the derivative of the objective with respect to the
// same-named forward plane; an adjoint state set has the State layout's
// first 11 planes (gas 5, face flux 3, cell B 3).  The MHD flux and EMF
// transposes are split at their linear stencils: a per-face (per-edge)
// pass writes stencil scalars to scratch planes and a per-cell (per-face)
// gather sums them, with no atomics.  The other kernels whose forward
// gathers over a stencil (the hydrodynamic flux, cell B, the boundary
// scatter of the inflow and outer-wall BCs) scatter with FP64 atomic adds
// (hardware global_atomic_add_f64); the rest gather.  dt is the forward
// run's value at each step; with the dt sensitivities on, the update, EMF
// and predictor transposes also write each point's share of dJ/d(dt) of
// their step into two planes (pdt: terms in the step's dt; pw: terms in
// the predictor's AB2 weight), which k_dt_part / k_dt_fold reduce into
// one value per step.  CalcDT's minimum (step.F90:58-101) is recorded in
// the forward steps (k_dtmin_part / k_dtmin_record: the argmin cell and its
// values before the boundary pass) and k_dt_inject carries the next step's
// value back through it.
#include "gam_adj.h"
#include "gam_kernels.h"

#define ABX 64
#define ABY 4
#define ABT 256   // threads of a point-kernel workgroup (one 256-point chunk)

// Point kernels: the points of each (i, j) plane of a box, i fastest, in
// 256-point chunks (no idle lanes whatever the row length), k from
// blockIdx.y.  Chunks are numbered so that workgroup w runs chunk
// (w % 8) * per + w / 8: the hardware puts workgroup w on XCD w % 8, so
// each XCD gets one contiguous run of chunks (which share stencil rows) in
// its own L2.
static inline dim3 agrid(const Box &b)
{
    const long nij = (long)(b.i1 - b.i0 + 1) * (b.j1 - b.j0 + 1);
    const unsigned nc = (unsigned)((nij + ABT - 1) / ABT);
    return dim3(((nc + 7) / 8) * 8, (unsigned)(b.k1 - b.k0 + 1));
}

#define POINT_IN(b)                                                                          \
    const int ni_ = (b).i1 - (b).i0 + 1;                                                     \
    const unsigned nij_ = (unsigned)ni_ * (unsigned)((b).j1 - (b).j0 + 1);                   \
    const unsigned L_ = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);            \
    const unsigned p_ = L_ * ABT + threadIdx.x;                                              \
    if (p_ >= nij_) return;                                                                  \
    const int j = (b).j0 + (int)(p_ / (unsigned)ni_);                                        \
    const int i = (b).i0 + (int)(p_ - (unsigned)(j - (b).j0) * (unsigned)ni_);              \
    const int k = (b).k0 + (int)blockIdx.y;

// Point kernels whose stencils run along j and k: the box's rows (the
// points of one (j, k) along i) in slabs of RBA j-rows, j fastest within a
// slab, then k, then the next slab, cut into 256-point chunks numbered as
// in POINT_IN (one contiguous run per XCD).  The rows at j +- 4 and at
// k +- 4 of a point are then at most 4 RBA rows away in the order the
// workgroups run, so the 8-point stencils of neighbouring workgroups meet
// in L2; with the plane-by-plane order of POINT_IN the k neighbours lie a
// whole (i, j) plane of workgroups away.
#define RBA 8
static inline dim3 agrid_slab(const Box &b)
{
    const long pts = (long)(b.i1 - b.i0 + 1) * (b.j1 - b.j0 + 1) * (b.k1 - b.k0 + 1);
    const long nc = (pts + ABT - 1) / ABT;
    return dim3((unsigned)(((nc + 7) / 8) * 8));
}

__device__ __forceinline__ void slab_row(const Box &b, unsigned nj, unsigned R, int &j, int &k)
{
    const unsigned nk = (unsigned)(b.k1 - b.k0 + 1);
    const unsigned slab = RBA * nk, nfull = nj / RBA, sb = R / slab;
    if (sb < nfull) {
        const unsigned r = R - sb * slab;
        k = b.k0 + (int)(r / RBA);
        j = b.j0 + (int)(sb * RBA + r % RBA);
    } else {
        const unsigned r = R - nfull * slab, rj = nj - nfull * RBA;
        k = b.k0 + (int)(r / rj);
        j = b.j0 + (int)(nfull * RBA + r % rj);
    }
}

#define POINT_SLAB(b)                                                                            \
    const unsigned ni_ = (unsigned)((b).i1 - (b).i0 + 1), nj_ = (unsigned)((b).j1 - (b).j0 + 1); \
    const unsigned L_ = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);                \
    const unsigned p_ = L_ * ABT + threadIdx.x;                                                  \
    if (p_ >= ni_ * nj_ * (unsigned)((b).k1 - (b).k0 + 1)) return;                               \
    const unsigned R_ = p_ / ni_;                                                                \
    const int i = (b).i0 + (int)(p_ - R_ * ni_);                                                 \
    int j, k;                                                                                    \
    slab_row((b), nj_, R_, j, k);

// the objective's block-sum reduction keeps plain 64 x 4 blocks
static inline dim3 agrid2(const Box &b)
{
    return dim3((unsigned)((b.i1 - b.i0 + ABX) / ABX), (unsigned)((b.j1 - b.j0 + ABY) / ABY),
                (unsigned)(b.k1 - b.k0 + 1));
}

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
// With pdt, active cells also write d/d(dtM) of their update: aSn . dH.
__global__ __launch_bounds__(256) void k_update_adj(Geo g, Box fb, Box ab, const Clock *clk,
                                                    const double *__restrict__ aSn,
                                                    double *__restrict__ aSo,
                                                    double *__restrict__ ag,
                                                    const double *__restrict__ met,
                                                    const double *__restrict__ gflx,
                                                    double *__restrict__ pdt)
{
    POINT_IN(fb);
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    if (i <= g.ieg && j <= g.jeg && k <= g.keg)
        for (int v = 0; v < 5; v++) aSo[v * N3 + c] += aSn[v * N3 + c];
    if (pdt && inside(ab, i, j, k)) {
        const double idV = 1.0 / met[M_VOL * N3 + c];
        double sdt = 0.0;
        for (int v = 0; v < 5; v++) {
            const double *gv = gflx + (long)v * N3;
            const double dH = (gv[c] - gv[c + 1] + gv[5 * N3 + c] - gv[5 * N3 + c + g.P] + gv[10 * N3 + c] -
                               gv[10 * N3 + c + g.PP]) * idV;
            sdt += aSn[v * N3 + c] * dH;
        }
        pdt[c] = sdt;
    }
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
    // the stencil values are loaded again (cached) for the transpose below
    // rather than kept live through the flux adjoint
    double WL[5], WR[5];
    for (int v = 0; v < 5; v++) {
        double qv[8], nq[4];
        for (int n = 0; n < 8; n++) qv[n] = H[(H_WH + v) * N3 + c + (n - 4) * s];
        for (int n = 0; n < 4; n++) nq[n] = H[(H_WN + v) * N3 + c + (n - 2) * s];
        recon_lr_w<RM>(m.pdmb, rVi, nq, qv, WL[v], WR[v]);
    }
    double T[9];
    for (int q = 0; q < 9; q++) T[q] = met[(M_TF + q + 9 * D) * N3 + c];
    const double faB = met[(M_FACE + D) * N3 + c];
    double aF[5], aL[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, aR[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int v = 0; v < 5; v++) aF[v] = faB * ag[(v + 5 * D) * N3 + c];
    gas_kin_adj(m.igm1, -1.0, WL, T, aF, aL);
    gas_kin_adj(m.igm1, +1.0, WR, T, aF, aR);
    for (int v = 0; v < 5; v++) {
        double qv[8], nq[4], aqv[8] = {0, 0, 0, 0, 0, 0, 0, 0}, anq[4] = {0, 0, 0, 0};
        for (int n = 0; n < 8; n++) qv[n] = H[(H_WH + v) * N3 + c + (n - 4) * s];
        for (int n = 0; n < 4; n++) nq[n] = H[(H_WN + v) * N3 + c + (n - 2) * s];
        recon_lr_w_adj<RM>(m.pdmb, rVi, nq, qv, aL[v], aR[v], anq, aqv);
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
                                                     double *__restrict__ aO,
                                                     double *__restrict__ pw)
{
    POINT_IN(cb);
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    const double w = ab2_weight(clk);
    double Q[5], oQ[5], aHG[5] = {0, 0, 0, 0, 0}, awh[5], awn[5], aQ[5] = {0, 0, 0, 0, 0},
           aoQ[5] = {0, 0, 0, 0, 0}, aw = 0.0;
    for (int v = 0; v < 5; v++) {
        Q[v] = S[(S_GAS + v) * N3 + c];
        oQ[v] = O[(S_GAS + v) * N3 + c];
        awh[v] = aWH[v * N3 + c];
        awn[v] = aWN[v * N3 + c];
    }
    predict_cell_adj(m, Q, oQ, w, met[M_VOL * N3 + c], aHG, awh, awn, aQ, aoQ, pw ? &aw : nullptr);
    if (pw) pw[c] = aw;
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

// dt = dtA = dts[n] (replays with a given dt sequence)
__global__ void k_set_dt_val(Clock *clk, const double *dts, int n)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Clock c = *clk;
    c.dt = dts[n];
    c.dtA = dts[n];
    *clk = c;
}

// ------------------------------------------------------------ dt sensitivities
// d(objective)/d(dt) of a step.  Its dt enters the update (dt dH, dt dM and
// the CT term dt curl E), the EMF's vDiff cap (CFL edge / dt) and the
// predictor's AB2 weight w = dt / (2 odt), with odt = State%time -
// oState%time, the previous step's dt.  The transposes leave each point's
// share in two planes: pdt (terms in the step's dt) and pw (terms in w).
// k_dt_part sums them per 64 x 4 block of the face box; k_dt_fold (one
// block, a fixed order: deterministic) folds the sums into mu.
__global__ __launch_bounds__(256) void k_dt_part(Geo g, Box b, const double *__restrict__ pdt,
                                                 const double *__restrict__ pw, double *__restrict__ part)
{
    __shared__ double r0[ABX * ABY], r1[ABX * ABY];
    const int tid = threadIdx.x + ABX * threadIdx.y;
    const int i = b.i0 + (int)(blockIdx.x * ABX + threadIdx.x);
    const int j = b.j0 + (int)(blockIdx.y * ABY + threadIdx.y);
    const int k = b.k0 + (int)blockIdx.z;
    double x = 0.0, y = 0.0;
    if (i <= b.i1 && j <= b.j1) {
        const long c = IX(g, i, j, k);
        x = pdt[c];
        y = pw[c];
    }
    r0[tid] = x;
    r1[tid] = y;
    __syncthreads();
    for (int h = ABX * ABY / 2; h > 0; h >>= 1) {
        if (tid < h) {
            r0[tid] += r0[tid + h];
            r1[tid] += r1[tid + h];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const long q = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
        part[2 * q] = r0[0];
        part[2 * q + 1] = r1[0];
    }
}

// mu[n] += the step's direct terms + the weight's share through dtA;
// mu[n - 1] += the weight's share through odt.  On a first step (oState%time
// = State%time) odt is the step's own dt, as in k_predict, and w = 1/2.
__global__ __launch_bounds__(256) void k_dt_fold(const double *__restrict__ part, long np,
                                                 const Clock *clk, int n, double *__restrict__ mu)
{
    __shared__ double r0[256], r1[256];
    const int tid = threadIdx.x;
    double x = 0.0, y = 0.0;
    for (long q = tid; q < np; q += 256) {
        x += part[2 * q];
        y += part[2 * q + 1];
    }
    r0[tid] = x;
    r1[tid] = y;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) {
            r0[tid] += r0[tid + h];
            r1[tid] += r1[tid + h];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const Clock c = *clk;
        const double st = c.st;
        double ost = c.ost;
        const bool first = fabs(ost - st) < GAM_TINY;
        if (first) ost = st - c.dt;
        const double odt = st - ost;
        const double w = (0.5 * c.dtA) / odt;
        const double aw = r1[0];
        double m = r0[0] + aw * (0.5 / odt);   // dw/d(dtA) = 1 / (2 odt)
        if (first) m += aw * (-w / odt);       // dw/d(odt) = -w / odt
        mu[n] += m;
        if (!first && n > 0) mu[n - 1] += aw * (-w / odt);
    }
}

// CalcDT's minimum (step.F90:58-101) over the DT box of the state the
// update has just written, before the boundary pass (run inside the
// recorded and recomputed forward steps, where that state exists): per
// 64 x 4 block the smallest CellDT and its cell (the lowest index on ties),
// keeping only positive finite values as the forward's reduction does.
__device__ __forceinline__ void dt_cell_in(const Geo &g, const MParm &m, long c, const double *S, const double *met,
                                           const double *bg, double *U, double *B, double &dl)
{
    const long N3 = g.N3;
    for (int v = 0; v < 5; v++) U[v] = S[(S_GAS + v) * N3 + c];
    for (int x = 0; x < 3; x++) B[x] = m.doMHD ? S[(S_B + x) * N3 + c] + (m.doBG ? bg[(B_B0 + x) * N3 + c] : 0.0) : 0.0;
    dl = dmin(dmin(met[M_DI * N3 + c], met[M_DJ * N3 + c]), met[M_DK * N3 + c]);
}

__device__ __forceinline__ void min_pair(double *rv, double *ri, int a, int b)
{
    if (ri[b] >= 0.0 && (ri[a] < 0.0 || rv[b] < rv[a] || (rv[b] == rv[a] && ri[b] < ri[a]))) {
        rv[a] = rv[b];
        ri[a] = ri[b];
    }
}

template <bool BORIS>
__global__ __launch_bounds__(256) void k_dtmin_part(Geo g, MParm m, Box b, const double *__restrict__ S,
                                                    const double *__restrict__ met,
                                                    const double *__restrict__ bg, double *__restrict__ part)
{
    __shared__ double rv[ABX * ABY], ri[ABX * ABY];
    const int tid = threadIdx.x + ABX * threadIdx.y;
    const int i = b.i0 + (int)(blockIdx.x * ABX + threadIdx.x);
    const int j = b.j0 + (int)(blockIdx.y * ABY + threadIdx.y);
    const int k = b.k0 + (int)blockIdx.z;
    double v = GAM_HUGE, ix = -1.0;
    if (i <= b.i1 && j <= b.j1) {
        const long c = IX(g, i, j, k);
        double U[5], B[3], dl;
        dt_cell_in(g, m, c, S, met, bg, U, B, dl);
        const double d = cell_dt<BORIS>(m, U, B[0], B[1], B[2], dl);
        if (d > 0.0 && d < GAM_HUGE) {
            v = d;
            ix = (double)c;
        }
    }
    rv[tid] = v;
    ri[tid] = ix;
    __syncthreads();
    for (int h = ABX * ABY / 2; h > 0; h >>= 1) {
        if (tid < h) min_pair(rv, ri, tid, tid + h);
        __syncthreads();
    }
    if (tid == 0) {
        const long q = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
        part[2 * q] = rv[0];
        part[2 * q + 1] = ri[0];
    }
}

// One block: the argmin over the blocks.  If its CellDT is the dt CalcDT
// has just set (clk, after the step's time bookkeeping; not capped by tFin
// or 10 dt0, where dt does not follow the state), rec = {cell, gas (5),
// total cell B (3), dl, 1}; otherwise rec[10] = 0.
template <bool BORIS>
__global__ __launch_bounds__(256) void k_dtmin_record(Geo g, MParm m, const double *__restrict__ part, long np,
                                                      const Clock *clk, const double *__restrict__ S,
                                                      const double *__restrict__ met,
                                                      const double *__restrict__ bg, double *__restrict__ rec)
{
    __shared__ double rv[256], ri[256];
    const int tid = threadIdx.x;
    rv[tid] = GAM_HUGE;
    ri[tid] = -1.0;
    for (long q = tid; q < np; q += 256) {
        const double pv = part[2 * q], pi = part[2 * q + 1];
        if (pi >= 0.0 && (ri[tid] < 0.0 || pv < rv[tid] || (pv == rv[tid] && pi < ri[tid]))) {
            rv[tid] = pv;
            ri[tid] = pi;
        }
    }
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) min_pair(rv, ri, tid, tid + h);
        __syncthreads();
    }
    if (tid != 0) return;
    const double dtn = clk->dt;
    rec[10] = 0.0;
    if (ri[0] < 0.0 || !(fabs(rv[0] - dtn) <= 1e-12 * dtn)) return;
    const long c = (long)ri[0];
    double U[5], B[3], dl;
    dt_cell_in(g, m, c, S, met, bg, U, B, dl);
    rec[0] = (double)c;
    for (int v = 0; v < 5; v++) rec[1 + v] = U[v];
    for (int x = 0; x < 3; x++) rec[6 + x] = B[x];
    rec[9] = dl;
    rec[10] = 1.0;
}

// CalcDT transpose: lam += mu * d(CellDT)/d(gas, cell B) at the recorded
// argmin (the adjoint of the update's output before the boundary pass).
template <bool BORIS>
__global__ void k_dt_inject(Geo g, MParm m, const double *__restrict__ rec, const double *__restrict__ mu,
                            double *__restrict__ lam)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double a = *mu;
    if (rec[10] == 0.0 || a == 0.0) return;
    const long c = (long)rec[0];
    double aU[5] = {0, 0, 0, 0, 0}, aB[3] = {0, 0, 0};
    cell_dt_adj<BORIS>(m, rec + 1, rec[6], rec[7], rec[8], rec[9], a, aU, aB);
    const long N3 = g.N3;
    for (int v = 0; v < 5; v++) lam[(S_GAS + v) * N3 + c] += aU[v];
    if (m.doMHD)
        for (int x = 0; x < 3; x++) lam[(S_B + x) * N3 + c] += aB[x];
}

// ============================================================ MHD
#define AADD(ptr, v) unsafeAtomicAdd((ptr), (v))

// Transpose of bFlux2Fld over cells (ghost cells, or every cell with
// all = true): each cell's B adjoint goes to its six face fluxes.
__global__ __launch_bounds__(256) void k_cellb_adj(Geo g, Box cb, int all, double *__restrict__ aS,
                                                   const double *__restrict__ met)
{
    POINT_IN(cb);
    if (!all && i >= g.is && i <= g.ie && j >= g.js && j <= g.je && k >= g.ks && k <= g.ke) return;
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    double aB[3], af[6] = {0, 0, 0, 0, 0, 0};
    for (int x = 0; x < 3; x++) {
        aB[x] = aS[(S_B + x) * N3 + c];
        aS[(S_B + x) * N3 + c] = 0.0;
    }
    cell_b_adj(g, met, c, aB, af);
    double *mf = aS + S_MF * N3;
    AADD(mf + c, af[0]);
    AADD(mf + N3 + c, af[1]);
    AADD(mf + 2 * N3 + c, af[2]);
    AADD(mf + c + 1, af[3]);
    AADD(mf + N3 + c + g.P, af[4]);
    AADD(mf + 2 * N3 + c + g.PP, af[5]);
}

__device__ __forceinline__ void move_adj(double *p, long dst, long src)
{
    p[src] += p[dst];
    p[dst] = 0.0;
}

// Transpose of the periodic BC's copies (k_bc_periodic): gas and cell B of
// ghost cells, and the face fluxes with the outer sides' shifted faces.
// Block z selects the planes: 0 the face fluxes, 1..5 one gas plane, 6..8
// one cell-B plane (each plane's chain of moves is independent of the
// others', so they run in parallel instead of one after another).
template <int SIDE>
__global__ void k_bc_periodic_adj(Geo g, Box t, double *__restrict__ aS)
{
    const int a = t.i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int b = t.j0 + (int)blockIdx.y;
    if (a > t.i1 || b > t.j1) return;
    const long N3 = g.N3;
    const int z = (int)blockIdx.z;
    const int pl = z <= 5 ? S_GAS + z - 1 : S_B + z - 6;
    double *mf = aS + S_MF * N3;
    for (int n = 4; n >= 1; n--) {
        int gi, gj, gk, si, sj, sk;
        if (SIDE == 0) { gi = g.is - n; si = g.ie - n + 1; gj = sj = a; gk = sk = b; }
        else if (SIDE == 1) { gi = g.ie + n; si = g.is + n - 1; gj = sj = a; gk = sk = b; }
        else if (SIDE == 2) { gj = g.js - n; sj = g.je - n + 1; gi = si = a; gk = sk = b; }
        else if (SIDE == 3) { gj = g.je + n; sj = g.js + n - 1; gi = si = a; gk = sk = b; }
        else if (SIDE == 4) { gk = g.ks - n; sk = g.ke - n + 1; gi = si = a; gj = sj = b; }
        else { gk = g.ke + n; sk = g.ks + n - 1; gi = si = a; gj = sj = b; }
        const long d = IX(g, gi, gj, gk), s = IX(g, si, sj, sk);
        // face fluxes (reverse of the forward statement order)
        if (z > 0) {
        } else if (SIDE == 1) {
            move_adj(mf, IX(g, gi + 1, gj, gk), IX(g, g.is + n, gj, gk));
            move_adj(mf + 2 * N3, d, s);
            move_adj(mf + N3, d, s);
        } else if (SIDE == 3) {
            move_adj(mf + N3, IX(g, gi, gj + 1, gk), IX(g, gi, g.js + n, gk));
            move_adj(mf + 2 * N3, d, s);
            move_adj(mf, d, s);
        } else if (SIDE == 5) {
            move_adj(mf + 2 * N3, IX(g, gi, gj, gk + 1), IX(g, gi, gj, g.ks + n));
            move_adj(mf + N3, d, s);
            move_adj(mf, d, s);
        } else {
            for (int x = 2; x >= 0; x--) move_adj(mf + x * N3, d, s);
        }
        if (z > 0 && gi <= g.ieg && gj <= g.jeg && gk <= g.keg) move_adj(aS + pl * N3, d, s);
    }
}

// Transpose of the zero-gradient / sphere BC copies with the magnetic
// face fluxes (k_bc_copy<SIDE, KIND>, KIND 1 zero gradient, 2 sphere in,
// 3 sphere out).  One thread per transverse point, all four layers; block
// z selects the planes as in k_bc_periodic_adj (the cell-B planes only move
// with the sphere copies).
template <int SIDE, int KIND>
__global__ void k_bc_copy_adj(Geo g, Box t, double *__restrict__ aS, const double *__restrict__ met)
{
    const int a = t.i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int b = t.j0 + (int)blockIdx.y;
    if (a > t.i1 || b > t.j1) return;
    const long N3 = g.N3;
    const int z = (int)blockIdx.z;
    const bool movesB = (SIDE == 0 && KIND == 2) || (SIDE == 1 && KIND == 3);
    if (z > 5 && !movesB) return;
    if (z > 0) {
        // the cell planes: plane pl of every ghost layer into the source cell
        const int pl = z <= 5 ? S_GAS + z - 1 : S_B + z - 6;
        double *p = aS + pl * N3;
        for (int n = 4; n >= 1; n--) {
            long d, sc;
            if (SIDE == 0) { d = IX(g, g.is - n, a, b); sc = IX(g, g.is, a, b); }
            else if (SIDE == 1) { d = IX(g, g.ie + n, a, b); sc = IX(g, g.ie, a, b); }
            else if (SIDE == 2) { d = IX(g, a, g.js - n, b); sc = IX(g, a, g.js, b); }
            else if (SIDE == 3) { d = IX(g, a, g.je + n, b); sc = IX(g, a, g.je, b); }
            else if (SIDE == 4) { d = IX(g, a, b, g.ks - n); sc = IX(g, a, b, g.ks); }
            else { d = IX(g, a, b, g.ke + n); sc = IX(g, a, b, g.ke); }
            move_adj(p, d, sc);
        }
        return;
    }
    // z = 0: the face fluxes
    double *mf = aS + S_MF * N3;
    if (SIDE == 0) {
        const long s = IX(g, g.is, a, b);
        for (int n = 4; n >= 1; n--) {
            const long d = IX(g, g.is - n, a, b);
            for (int x = 2; x >= 0; x--) move_adj(mf + x * N3, d, s);
        }
    } else if (SIDE == 1) {
        const long s = IX(g, g.ie, a, b);
        if (KIND == 3) {
            for (int n = 4; n >= 1; n--) {
                const int ig = g.ie + n;
                const long d = IX(g, ig, a, b);
                move_adj(mf + 2 * N3, d, s);
                move_adj(mf + N3, d, s);
                move_adj(mf, IX(g, ig + 1, a, b), s);
            }
        } else {
            // ghost fluxes = face area x (B . normal), B = cell B of the last
            // active cell from its own six face fluxes
            double aB[3] = {0.0, 0.0, 0.0};
            for (int n = 4; n >= 1; n--) {
                const int ig = g.ie + n;
                const long d = IX(g, ig, a, b);
                const long fi[3] = {IX(g, ig + 1, a, b), d, d};
                for (int x = 2; x >= 0; x--) {
                    const long f = fi[x];
                    const double am = mf[x * N3 + f];
                    mf[x * N3 + f] = 0.0;
                    const double A = met[(M_FACE + x) * N3 + f];
                    for (int y = 0; y < 3; y++) aB[y] += A * met[(M_TF + 9 * x + y) * N3 + f] * am;
                }
            }
            // the J and K faces of the last cell are shared with the
            // neighbouring transverse threads: atomic adds
            double af[6] = {0, 0, 0, 0, 0, 0};
            cell_b_adj(g, met, s, aB, af);
            AADD(mf + s, af[0]);
            AADD(mf + N3 + s, af[1]);
            AADD(mf + 2 * N3 + s, af[2]);
            AADD(mf + s + 1, af[3]);
            AADD(mf + N3 + s + g.P, af[4]);
            AADD(mf + 2 * N3 + s + g.PP, af[5]);
        }
    } else if (SIDE == 2) {
        const long s = IX(g, a, g.js, b);
        for (int n = 4; n >= 1; n--) {
            const long d = IX(g, a, g.js - n, b);
            for (int x = 2; x >= 0; x--) move_adj(mf + x * N3, d, s);
        }
    } else if (SIDE == 3) {
        const long s = IX(g, a, g.je, b);
        for (int n = 4; n >= 1; n--) {
            const long d = IX(g, a, g.je + n, b);
            move_adj(mf + N3, IX(g, a, g.je + n + 1, b), IX(g, a, g.je + 1, b));
            move_adj(mf + 2 * N3, d, s);
            move_adj(mf, d, s);
        }
    } else if (SIDE == 4) {
        const long s = IX(g, a, b, g.ks);
        for (int n = 4; n >= 1; n--) {
            const long d = IX(g, a, b, g.ks - n);
            for (int x = 2; x >= 0; x--) move_adj(mf + x * N3, d, s);
        }
    } else {
        const long s = IX(g, a, b, g.ke);
        for (int n = 4; n >= 1; n--) {
            const long d = IX(g, a, b, g.ke + n);
            // the k+1 ghost face takes the J flux (bcs.F90:925)
            const long fk = IX(g, a, b, g.ke + n + 1), fj = IX(g, a, b, g.ke + 1);
            mf[N3 + fj] += mf[2 * N3 + fk];
            mf[2 * N3 + fk] = 0.0;
            move_adj(mf + N3, d, s);
            move_adj(mf, d, s);
        }
    }
}

// Update transpose, part 1 (cells): gas of the Reynolds + Maxwell/Boris
// update (active cells) or the copy (other cells).  Writes adH (adjoint of
// the gas flux divergence) and adM (of the magnetic stress divergence);
// with Boris also adds the adjoints of the new and old cell B (total B
// minus the constant background) into aSn and aSo.  Sn1 = the step's
// output State (its cell B is the new B of the update).
template <bool BORIS>
__global__ __launch_bounds__(256) void k_update_adj_cells(Geo g, MParm m, Box cb, Box ab,
                                                          const Clock *clk,
                                                          const double *__restrict__ So,
                                                          const double *__restrict__ Sn1,
                                                          const double *__restrict__ gflx,
                                                          const double *__restrict__ mflx,
                                                          const double *__restrict__ met,
                                                          const double *__restrict__ bg, int BG,
                                                          double *__restrict__ aSn,
                                                          double *__restrict__ aSo,
                                                          double *__restrict__ adHo,
                                                          double *__restrict__ adMo,
                                                          double *__restrict__ pdt)
{
    POINT_IN(cb);
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    if (!inside(ab, i, j, k)) {
        for (int v = 0; v < 5; v++) aSo[v * N3 + c] += aSn[v * N3 + c];
        return;
    }
    const double idV = 1.0 / met[M_VOL * N3 + c];
    const double dtM = clk->dt, dt = clk->dtA;
    const long oI = c + 1, oJ = c + g.P, oK = c + g.PP;
    double oU[5], dH[5], dM[3], aUp[5], aoU[5] = {0, 0, 0, 0, 0}, adH[5] = {0, 0, 0, 0, 0},
           adm[3] = {0, 0, 0};
    for (int v = 0; v < 5; v++) {
        const double *gv = gflx + (long)v * N3;
        dH[v] = (gv[c] - gv[oI] + gv[5 * N3 + c] - gv[5 * N3 + oJ] + gv[10 * N3 + c] - gv[10 * N3 + oK]) * idV;
        oU[v] = So[v * N3 + c];
        aUp[v] = aSn[v * N3 + c];
    }
    for (int x = 0; x < 3; x++) {
        const double *mv = mflx + (long)x * N3;
        double r = (mv[c] - mv[oI] + mv[3 * N3 + c] - mv[3 * N3 + oJ] + mv[6 * N3 + c] - mv[6 * N3 + oK]) * idV;
        if (BG) r = r + bg[(B_DPB0 + x) * N3 + c];
        dM[x] = r;
    }
    if (BORIS) {
        double B[3], oB[3], aB[3] = {0, 0, 0}, aoB[3] = {0, 0, 0};
        for (int x = 0; x < 3; x++) {
            const double b0 = BG ? bg[(B_B0 + x) * N3 + c] : 0.0;
            B[x] = Sn1[(S_B + x) * N3 + c] + b0;
            oB[x] = So[(S_B + x) * N3 + c] + b0;
        }
        boris_adj(m, oU, dH, dM, B, oB, dtM, dt, aUp, aoU, adH, adm, aB, aoB);
        for (int x = 0; x < 3; x++) {
            aSn[(S_B + x) * N3 + c] += aB[x];
            aSo[(S_B + x) * N3 + c] += aoB[x];
        }
    } else {
        double U[5], aU[5] = {0, 0, 0, 0, 0};
        for (int v = 0; v < 5; v++) U[v] = oU[v] + dtM * dH[v];
        maxwell_adj(m, U, dM, dt, aUp, aU, adm);
        for (int v = 0; v < 5; v++) {
            aoU[v] = aU[v];
            adH[v] = dtM * aU[v];
        }
    }
    for (int v = 0; v < 5; v++) {
        aSo[v * N3 + c] += aoU[v];
        adHo[v * N3 + c] = adH[v];
    }
    for (int x = 0; x < 3; x++) adMo[x * N3 + c] = adm[x];
    if (pdt) {
        // the cell's update depends on dt only through dt dH and dt dM, so
        // dt d/d(dt) = dH . adH + dM . adM
        double sdt = 0.0;
        for (int v = 0; v < 5; v++) sdt += dH[v] * adH[v];
        for (int x = 0; x < 3; x++) sdt += dM[x] * adm[x];
        pdt[c] = sdt / dt;
    }
}

// Update transpose, part 2 (faces): adjoints of the active faces' gas and
// magnetic fluxes, of the old face fluxes (copy + CT) and of the edge EMFs.
__global__ __launch_bounds__(256) void k_update_adj_faces(Geo g, Box fb, Box ab, Box ctb, EdgeBoxes eb,
                                                          const Clock *clk,
                                                          const double *__restrict__ met,
                                                          const double *__restrict__ aUr,
                                                          const double *__restrict__ adM,
                                                          const double *__restrict__ aSn,
                                                          double *__restrict__ aSo,
                                                          double *__restrict__ ag,
                                                          double *__restrict__ amf,
                                                          double *__restrict__ aE)
{
    POINT_IN(fb);
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    const double dtA = clk->dtA;
    const long st[3] = {1, g.P, g.PP};
    for (int d = 0; d < 3; d++) {
        const bool face = i >= g.is && j >= g.js && k >= g.ks && i <= g.ie + (d == 0) &&
                          j <= g.je + (d == 1) && k <= g.ke + (d == 2);
        if (!face) continue;
        const long cl = c - st[d];
        const bool up = inside(ab, i, j, k);
        const bool lo = inside(ab, i - (d == 0), j - (d == 1), k - (d == 2));
        const double wu = up ? 1.0 / met[M_VOL * N3 + c] : 0.0;
        const double wl = lo ? 1.0 / met[M_VOL * N3 + cl] : 0.0;
        // adH (in aUr), adM hold values at active cells only
        for (int v = 0; v < 5; v++)
            ag[(v + 5 * d) * N3 + c] = (up ? wu * aUr[v * N3 + c] : 0.0) - (lo ? wl * aUr[v * N3 + cl] : 0.0);
        for (int x = 0; x < 3; x++)
            amf[(x + 3 * d) * N3 + c] = (up ? wu * adM[x * N3 + c] : 0.0) - (lo ? wl * adM[x * N3 + cl] : 0.0);
    }
    const double *aF = aSn + S_MF * N3;   // adjoint of the new face fluxes
    for (int d = 0; d < 3; d++) aSo[(S_MF + d) * N3 + c] += aF[d * N3 + c];
    // CT: new flux(d, f) = old - dtA * (curl of E around f); gather per edge
    auto fa = [&](int d, int ii, int jj, int kk) -> double {
        return inside(ctb, ii, jj, kk) ? aF[d * N3 + IX(g, ii, jj, kk)] : 0.0;
    };
    aE[c] = inside(eb.e[0], i, j, k)
                ? dtA * (fa(1, i, j, k) - fa(1, i, j, k - 1) + fa(2, i, j - 1, k) - fa(2, i, j, k)) : 0.0;
    aE[N3 + c] = inside(eb.e[1], i, j, k)
                ? dtA * (fa(0, i, j, k - 1) - fa(0, i, j, k) + fa(2, i, j, k) - fa(2, i - 1, j, k)) : 0.0;
    aE[2 * N3 + c] = inside(eb.e[2], i, j, k)
                ? dtA * (fa(0, i, j, k) - fa(0, i, j - 1, k) + fa(1, i - 1, j, k) - fa(1, i, j, k)) : 0.0;
}

// Interface fluxes' transpose with the magnetic terms: gas and magnetic
// kinetic fluxes (background-field cross terms, Boris factor), HOGS, the
// Boris mag-HOGS and nuke terms, L/R states of gas (H_WH, H_WN) and cell B
// (H_B weighted by volume, limiter S_B), the normal flux Bn = H_MF / A.
template <int D, int RM, bool BG, bool BORIS>
__global__ __launch_bounds__(256, 2) void k_flux_adj_mhd(Geo g, MParm m, Box fb,
                                                      const double *__restrict__ H,
                                                      const double *__restrict__ S,
                                                      const double *__restrict__ met,
                                                      const double *__restrict__ bg,
                                                      const double *__restrict__ ag,
                                                      const double *__restrict__ amf,
                                                      double *__restrict__ sc,
                                                      double *__restrict__ aHMF,
                                                      double *__restrict__ aHG)
{
    POINT_SLAB(fb);
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    const long s = STRIDE<D>(g);
    // stencil point n of a plane: a 32-bit byte offset from the plane base
    // (as k_flux), formed where it is used rather than kept in registers
    // through the kernel; along i plain indexing lets the compiler merge
    // neighbouring points
    const unsigned oc = (unsigned)(c * 8), os = (unsigned)(s * 8);
    auto LD = [&](const double *__restrict__ plane, int n) -> double {
        if (D == 0) return plane[c + (n - 4)];
        return ldp(plane, oc + (unsigned)(n - 4) * os);
    };
    double vol[8];
    for (int n = 0; n < 8; n++) vol[n] = LD(met + M_VOL * N3, n);
    const double rVi = 1.0 / wsum8([&](int n) { return vol[n]; });
    double WL[5], WR[5], BL[3], BR[3];
    for (int v = 0; v < 5; v++) {
        double qv[8], nq[4];
        for (int n = 0; n < 8; n++) qv[n] = LD(H + (H_WH + v) * N3, n);
        for (int n = 0; n < 4; n++) nq[n] = LD(H + (H_WN + v) * N3, n + 2);
        recon_lr_w<RM>(m.pdmb, rVi, nq, qv, WL[v], WR[v]);
    }
    const double *Bl = m.doLFM ? S + S_B * N3 : H + H_B * N3;
    for (int x = 0; x < 3; x++) {
        double q[8], nq[4];
        for (int n = 0; n < 8; n++) q[n] = LD(H + (H_B + x) * N3, n);
        for (int n = 0; n < 4; n++) nq[n] = LD(Bl + x * N3, n + 2);
        recon_lr<RM>(m.pdmb, vol, rVi, nq, q, BL[x], BR[x]);
    }
    double T[9];
    for (int q = 0; q < 9; q++) T[q] = ldp(met + (M_TF + q + 9 * D) * N3, oc);
    const double faB = ldp(met + (M_FACE + D) * N3, oc);
    const double iFaB = 1.0 / faB;
    const double Bn = ldp(H + (H_MF + D) * N3, oc) * iFaB;
    double B0[3] = {0, 0, 0}, B0n = 0.0;
    if (BG) {
        for (int x = 0; x < 3; x++) B0[x] = ldp(bg + (B_FCB0 + x + 3 * D) * N3, oc);
        B0n = ldp(bg + (B_BFLUX0 + D) * N3, oc) * iFaB;
    }
    // forward pieces the adjoint needs: VaD, bbD and the HOGS differences dW
    double mF[3] = {0, 0, 0}, bbD = 0.0, VaD = 0.0;
    mag_kin<BG, BORIS>(m, -1.0, WL, BL, Bn, T, B0, B0n, mF, bbD, VaD);
    mag_kin<BG, BORIS>(m, +1.0, WR, BR, Bn, T, B0, B0n, mF, bbD, VaD);
    double gF[5] = {0, 0, 0, 0, 0}, dW[5] = {0, 0, 0, 0, 0};
    gas_kin(m.igm1, -1.0, WL, T, gF, dW);
    gas_kin(m.igm1, +1.0, WR, T, gF, dW);
    // adjoint
    double aF[5], adW[5], aM[3], aVaD = 0.0, abbD = 0.0;
    double aL[5] = {0, 0, 0, 0, 0}, aR[5] = {0, 0, 0, 0, 0}, abL[3] = {0, 0, 0}, abR[3] = {0, 0, 0}, aBn = 0.0;
    for (int v = 0; v < 5; v++) {
        aF[v] = faB * ldp(ag + (v + 5 * D) * N3, oc);
        adW[v] = -m.cHogH * VaD * aF[v];
        aVaD += -m.cHogH * dW[v] * aF[v];
    }
    for (int x = 0; x < 3; x++) aM[x] = faB * ldp(amf + (x + 3 * D) * N3, oc);
    if (BORIS) {
        // mag-HOGS: mF -= cHogM VaD bbD (WR - WL)_v / Ca^2
        for (int x = 0; x < 3; x++) {
            const double dv = WR[1 + x] - WL[1 + x];
            const double deeP = bbD * dv * m.iCa2;
            aVaD += -m.cHogM * deeP * aM[x];
            const double adeeP = -m.cHogM * VaD * aM[x];
            abbD += dv * m.iCa2 * adeeP;
            aR[1 + x] += bbD * m.iCa2 * adeeP;
            aL[1 + x] -= bbD * m.iCa2 * adeeP;
        }
        // nuke: gF -= Ca (U(c) - U(c - s)) where the mean speed is large
        const double vL = sqrt(WL[1] * WL[1] + WL[2] * WL[2] + WL[3] * WL[3]);
        const double vR = sqrt(WR[1] * WR[1] + WR[2] * WR[2] + WR[3] * WR[3]);
        const double vn = sqrt(vL * vL + vR * vR) / GAM_SQRT2_SP;
        if (vn >= 1.5 * m.Ca)
            for (int v = 0; v < 5; v++) {
                AADD(aHG + v * N3 + c, -m.Ca * aF[v]);
                AADD(aHG + v * N3 + c - s, m.Ca * aF[v]);
            }
    }
    gas_kin_adj(m.igm1, -1.0, WL, T, aF, aL, adW);
    gas_kin_adj(m.igm1, +1.0, WR, T, aF, aR, adW);
    mag_kin_adj<BG, BORIS>(m, -1.0, WL, BL, Bn, T, B0, B0n, aM, aVaD, abbD, aL, abL, aBn);
    mag_kin_adj<BG, BORIS>(m, +1.0, WR, BR, Bn, T, B0, B0n, aM, aVaD, abbD, aR, abR, aBn);
    AADD(aHMF + D * N3 + c, aBn * iFaB);
    // the reconstructions' transposes up to their linear stencils: per
    // variable (gas 5, then B 3) the stencil scalars sl, sr and the four
    // limiter adjoints, gathered into the cells by k_recon_adj_gather
    for (int v = 0; v < 5; v++) {
        double qv[8], nq[4], anq[4] = {0, 0, 0, 0}, sl, sr;
        for (int n = 0; n < 8; n++) qv[n] = LD(H + (H_WH + v) * N3, n);
        for (int n = 0; n < 4; n++) nq[n] = LD(H + (H_WN + v) * N3, n + 2);
        recon_lr_w_adj_s<RM>(m.pdmb, rVi, nq, qv, aL[v], aR[v], anq, sl, sr);
        *(double2 *)(sc + 6 * v * N3 + 2 * c) = make_double2(sl, sr);
        for (int n = 0; n < 4; n++) stp(sc + (6 * v + 2 + n) * N3, oc, anq[n]);
    }
    for (int x = 0; x < 3; x++) {
        double qv[8], nq[4], anq[4] = {0, 0, 0, 0}, sl, sr;
        // the volumes are loaded again (cached) rather than kept live through
        // the flux adjoint, which holds the kernel's register peak
        for (int n = 0; n < 8; n++) qv[n] = LD(met + M_VOL * N3, n) * LD(H + (H_B + x) * N3, n);
        for (int n = 0; n < 4; n++) nq[n] = LD(Bl + x * N3, n + 2);
        recon_lr_w_adj_s<RM>(m.pdmb, rVi, nq, qv, abL[x], abR[x], anq, sl, sr);
        const int v = 5 + x;
        *(double2 *)(sc + 6 * v * N3 + 2 * c) = make_double2(sl, sr);
        for (int n = 0; n < 4; n++) stp(sc + (6 * v + 2 + n) * N3, oc, anq[n]);
    }
}

// The cell side of the reconstruction transposes of the three face
// directions (k_flux_adj_mhd's scratch, 48 planes per direction) and of
// Mom2Face (rv, aVf): one thread per cell sums what the faces whose
// stencils hold it give it, with no atomics, and WRITES the half-state
// adjoints aWH (5), aWN (5) and aHB (3) of every cell; with the LFM
// limiter the B limiter adjoints go into aSSB (3, added).  Faces of
// direction D exist on the active box extended by one along D.
template <int RM>
__global__ __launch_bounds__(256) void k_recon_adj_gather(Geo g, Box cb, const double *__restrict__ sc,
                                                          const double *__restrict__ rv,
                                                          const double *__restrict__ aVf,
                                                          const double *__restrict__ met, int lfm,
                                                          double *__restrict__ aWH, double *__restrict__ aWN,
                                                          double *__restrict__ aHB, double *__restrict__ aSSB)
{
    POINT_SLAB(cb);
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    double gh[8], gn[8], av[3] = {0.0, 0.0, 0.0};
    for (int v = 0; v < 8; v++) gh[v] = gn[v] = 0.0;
    for (int D = 0; D < 3; D++) {
        const long s = D == 0 ? 1 : D == 1 ? (long)g.P : g.PP;
        const int at = D == 0 ? i : D == 1 ? j : k;
        const int lo = D == 0 ? g.is : D == 1 ? g.js : g.ks;
        const int hi = (D == 0 ? g.ie : D == 1 ? g.je : g.ke) + 1;
        // Mom2Face faces span every transverse cell
        for (int n = 0; n < 8; n++) {
            const int fi = at - (n - 4);
            if (fi < lo || fi > hi) continue;
            const long f = c - (n - 4) * s;
            const double w = W8(n) * rv[D * N3 + f];
            for (int x = 0; x < 3; x++) av[x] += w * aVf[(3 * D + x) * N3 + f];
        }
        // interface-flux faces: active transverse cells only
        const bool tr = D == 0 ? (j >= g.js && j <= g.je && k >= g.ks && k <= g.ke)
                      : D == 1 ? (i >= g.is && i <= g.ie && k >= g.ks && k <= g.ke)
                               : (i >= g.is && i <= g.ie && j >= g.js && j <= g.je);
        if (!tr) continue;
        const double *sd = sc + (long)48 * D * N3;
#pragma unroll
        for (int n = 0; n < 8; n++) {
            const int fi = at - (n - 4);   // this cell is stencil point n of face fi
            if (fi < lo || fi > hi) continue;
            const long f = c - (n - 4) * s;
            for (int v = 0; v < 8; v++) {
                const double2 lr = *(const double2 *)(sd + 6 * v * N3 + 2 * f);
                if (RM == 1) {
                    if (n <= 6) gh[v] += UP7(n) * lr.x;
                    if (n >= 1) gh[v] += UP7(7 - n) * lr.y;
                } else {
                    gh[v] += W8(n) * lr.x;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int fi = at - (q - 2);   // this cell is limiter point q of face fi
            if (fi < lo || fi > hi) continue;
            const long f = c - (q - 2) * s;
            for (int v = 0; v < 8; v++) gn[v] += sd[(6 * v + 2 + q) * N3 + f];
        }
    }
    for (int v = 0; v < 5; v++) {
        aWH[v * N3 + c] = gh[v] + ((v >= 1 && v <= 3) ? av[v - 1] : 0.0);
        aWN[v * N3 + c] = gn[v];
    }
    const double V = met[M_VOL * N3 + c];
    for (int x = 0; x < 3; x++) {
        aHB[x * N3 + c] = V * gh[5 + x] + (lfm ? 0.0 : gn[5 + x]);
        if (lfm) aSSB[x * N3 + c] += gn[5 + x];
    }
}

// Edge EMF transpose (k_emf), pass 1: per edge, the pieces of
// edge_emf_adj_s into sc (16 planes for this edge direction: sv 3, the
// pair (sl1, sr1) interleaved over two planes, anq1 4, the pair (sl2, sr2),
// anq2 4, aDc), gathered into the face velocities,
// the half-state and limiter face fluxes and the half-state density by
// k_emf_adj_gather.
template <int ED, int RM, bool BG, bool BORIS>
__global__ __launch_bounds__(256, 2) void k_emf_adj(Geo g, MParm m, Box eb, const Clock *clk,
                                                 const double *__restrict__ H,
                                                 const double *__restrict__ S,
                                                 const double *__restrict__ Vf,
                                                 const double *__restrict__ met,
                                                 const double *__restrict__ bg,
                                                 const double *__restrict__ aE,
                                                 double *__restrict__ sc,
                                                 double *__restrict__ pdt)
{
    POINT_SLAB(eb);
    constexpr int T1 = (ED == 0) ? 1 : (ED == 1) ? 2 : 0;
    constexpr int T2 = (ED == 0) ? 2 : (ED == 1) ? 0 : 1;
    constexpr int E1 = (ED == 0) ? 1 : 0;
    constexpr int E2 = (ED == 2) ? 1 : 2;
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    const long s1 = STRIDE<T1>(g), s2 = STRIDE<T2>(g);
    const long se1 = STRIDE<E1>(g), se2 = STRIDE<E2>(g);
    const double *mfH = H + H_MF * N3;
    const double *mfL = m.doLFM ? S + S_MF * N3 : mfH;
    EdgeIn in;
    for (int n = 0; n < 8; n++) {
        const long q1 = c + (n - 4) * s2, q2 = c + (n - 4) * s1;
        in.A1[n] = met[(M_FACE + T1) * N3 + q1];
        in.mH1[n] = mfH[T1 * N3 + q1];
        in.A2[n] = met[(M_FACE + T2) * N3 + q2];
        in.mH2[n] = mfH[T2 * N3 + q2];
    }
    for (int d = 0; d < 3; d++)
        in.AV[d] = wsum8([&](int n) { return in.A1[n] * Vf[(3 * T1 + d) * N3 + c + (n - 4) * s2]; });
    for (int n = 0; n < 4; n++) {
        in.mL1[n] = mfL[T1 * N3 + c + (n - 2) * s2];
        in.mL2[n] = mfL[T2 * N3 + c + (n - 2) * s1];
    }
    const double *Te = met + (long)(M_TE + 9 * ED) * N3 + c;
    for (int d = 0; d < 3; d++) { in.te1[d] = Te[(3 + d) * N3]; in.te2[d] = Te[(6 + d) * N3]; }
    const double *Tb = met + (long)(M_TEB + 4 * ED) * N3 + c;
    for (int q = 0; q < 4; q++) in.tb[q] = Tb[q * N3];
    const long dc[4] = {c, c - se1, c - se2, c - se1 - se2};
    for (int q = 0; q < 4; q++) in.D4[q] = H[(H_GAS + DEN) * N3 + dc[q]];
    in.b01 = BG ? bg[(B_EDGB0 + 0 + 2 * ED) * N3 + c] : 0.0;
    in.b02 = BG ? bg[(B_EDGB0 + 1 + 2 * ED) * N3 + c] : 0.0;
    in.edge = met[(M_EDGE + ED) * N3 + c];
    in.cap = m.CFL * in.edge / clk->dt;
    EdgeAdjS o;
    const double aEc = aE[ED * N3 + c];
    edge_emf_adj_s<RM, BORIS>(m, in, aEc, o);
    // dt of this edge's terms: constrained transport (new fluxes = old -
    // dt C E with C the circulation, and aE = -dt C^T aF, so the flux
    // adjoints' share is E . aE / dt) and the vDiff cap CFL edge / dt
    if (pdt) pdt[c] += (o.E * aEc - o.acap * in.cap) / clk->dt;
    const double v[16] = {o.sv[0], o.sv[1], o.sv[2], o.sl1, o.sr1, o.anq1[0], o.anq1[1], o.anq1[2], o.anq1[3],
                          o.sl2, o.sr2, o.anq2[0], o.anq2[1], o.anq2[2], o.anq2[3], o.aDc};
    for (int q = 0; q < 16; q++)
        if (q != 3 && q != 4 && q != 9 && q != 10) sc[q * N3 + c] = v[q];
    // the stencil scalar pairs interleaved over planes 3-4 and 9-10
    *(double2 *)(sc + 3 * N3 + 2 * c) = make_double2(o.sl1, o.sr1);
    *(double2 *)(sc + 9 * N3 + 2 * c) = make_double2(o.sl2, o.sr2);
}

// Edge EMF transpose, pass 2: one thread per face position x (every
// direction T) sums what the edges whose stencils hold it give it:
// the corner velocities of the edges of direction T-1 along T+1 (aVf),
// their edge-B reconstructions of the T fluxes and those of the edges of
// direction T+1 along T+2 (aHMF; the limiter parts added to aSSMF with
// the LFM limiter); and, as a cell, the corner density of the four edges
// of each direction around it (aHG's density; its other planes get 0).
// aVf, aHMF and aHG are WRITTEN at every point (the flux transpose adds
// into aHMF and aHG afterwards).  sc: 16 planes per edge direction from
// k_emf_adj; edges exist in eb.
template <int RM>
__global__ __launch_bounds__(256) void k_emf_adj_gather(Geo g, Box fb, EdgeBoxes eb, const double *__restrict__ sc,
                                                        const double *__restrict__ met, int lfm,
                                                        double *__restrict__ aVf, double *__restrict__ aHMF,
                                                        double *__restrict__ aHG, double *__restrict__ aSSMF)
{
    POINT_SLAB(fb);
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    const long st[3] = {1, (long)g.P, g.PP};
#pragma unroll
    for (int T = 0; T < 3; T++) {
        const int EDa = (T + 2) % 3, Da = (T + 1) % 3;   // edges with T1 = T, stencil along T2
        const int EDb = (T + 1) % 3, Db = (T + 2) % 3;   // edges with T2 = T, stencil along T1
        const double *sa = sc + (long)16 * EDa * N3, *sb = sc + (long)16 * EDb * N3;
        double sv[3] = {0.0, 0.0, 0.0}, sh = 0.0, sn = 0.0;
#pragma unroll
        for (int n = 0; n < 8; n++) {
            int e3[3] = {i, j, k};
            e3[Da] -= n - 4;
            if (inside(eb.e[EDa], e3[0], e3[1], e3[2])) {
                const long e = c - (n - 4) * st[Da];
                for (int d = 0; d < 3; d++) sv[d] += W8(n) * sa[d * N3 + e];
                const double2 lr = *(const double2 *)(sa + 3 * N3 + 2 * e);
                sh += stencil_adj<RM>(n, lr.x, lr.y);
            }
            int e4[3] = {i, j, k};
            e4[Db] -= n - 4;
            if (inside(eb.e[EDb], e4[0], e4[1], e4[2])) {
                const long e = c - (n - 4) * st[Db];
                const double2 lr = *(const double2 *)(sb + 9 * N3 + 2 * e);
                sh += stencil_adj<RM>(n, lr.x, lr.y);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            int e3[3] = {i, j, k};
            e3[Da] -= q - 2;
            if (inside(eb.e[EDa], e3[0], e3[1], e3[2])) sn += sa[(5 + q) * N3 + c - (q - 2) * st[Da]];
            int e4[3] = {i, j, k};
            e4[Db] -= q - 2;
            if (inside(eb.e[EDb], e4[0], e4[1], e4[2])) sn += sb[(11 + q) * N3 + c - (q - 2) * st[Db]];
        }
        const double A = met[(M_FACE + T) * N3 + c], Am = dmax(A, GAM_TINY);
        for (int d = 0; d < 3; d++) aVf[(3 * T + d) * N3 + c] = A * sv[d];
        aHMF[T * N3 + c] = (A / Am) * sh + (lfm ? 0.0 : sn / Am);
        if (lfm) aSSMF[T * N3 + c] += sn / Am;
    }
    if (i <= g.ieg && j <= g.jeg && k <= g.keg) {
        double sd = 0.0;
#pragma unroll
        for (int ED = 0; ED < 3; ED++) {
            const int E1 = (ED == 0) ? 1 : 0, E2 = (ED == 2) ? 1 : 2;
            for (int ab = 0; ab < 4; ab++) {
                int e3[3] = {i, j, k};
                e3[E1] += ab & 1;
                e3[E2] += ab >> 1;
                if (inside(eb.e[ED], e3[0], e3[1], e3[2]))
                    sd += sc[(16 * ED + 15) * N3 + c + (ab & 1) * st[E1] + (ab >> 1) * st[E2]];
            }
        }
        aHG[DEN * N3 + c] = 0.25 * sd;
        for (int v = 1; v < 5; v++) aHG[v * N3 + c] = 0.0;
    }
}

// Mom2Face (k_vface): face velocity of direction d at face f =
// sum_n W8(n) V_n v_n / sum_n W8(n) V_n over the 8 cells along d (cells
// past the k ends of the grid left out).  rv[d] = that normalisation,
// 1 / max(sum_n W8(n) V_n, tiny), at each face k_vface writes (0
// elsewhere); geometry only, formed once per adjoint run.
__global__ __launch_bounds__(256) void k_vface_rv(Geo g, Box cb, const double *__restrict__ met,
                                                  double *__restrict__ rv)
{
    POINT_IN(cb);
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    const double *vol = met + (long)M_VOL * N3;
    for (int d = 0; d < 3; d++) {
        const bool on = d == 0 ? (i >= g.is && i <= g.ie + 1)
                      : d == 1 ? (j >= g.js && j <= g.je + 1) : (k >= g.ks && k <= g.ke + 1);
        double r = 0.0;
        if (on) {
            const long s = d == 0 ? 1 : d == 1 ? (long)g.P : g.PP;
            double v8[8];
            for (int n = 0; n < 8; n++) {
                const int kk = k + (n - 4);
                v8[n] = (d != 2 || (kk >= g.ksg && kk <= g.keg)) ? vol[c + (n - 4) * s] : 0.0;
            }
            r = 1.0 / dmax(wsum8([&](int n) { return v8[n]; }), GAM_TINY);
        }
        rv[d * N3 + c] = r;
    }
}

// The predictor's face fluxes (AB2 of State and oState) and, gathered per
// face, the transpose of the half-state cell B it forms from them (cell_b
// of the cells on either side of the face: this face is the low face of
// cell c and the high face of cell c - s).
__global__ __launch_bounds__(256) void k_predict_mf_adj(Geo g, Box fb, const Clock *clk,
                                                        const double *__restrict__ met,
                                                        const double *__restrict__ aHB,
                                                        const double *__restrict__ aHMF,
                                                        double *__restrict__ aS,
                                                        double *__restrict__ aO,
                                                        const double *__restrict__ S,
                                                        const double *__restrict__ O,
                                                        double *__restrict__ pw)
{
    POINT_IN(fb);
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    const double w = ab2_weight(clk);
    const bool cin = i <= g.ieg && j <= g.jeg && k <= g.keg;   // cell c exists
    for (int d = 0; d < 3; d++) {
        double a = aHMF[d * N3 + c];
        const double *xf = met + (long)(M_XFC + 3 * d) * N3;
        if (cin) {
            const double iv = 1.0 / met[M_VOL * N3 + c];
            for (int x = 0; x < 3; x++)
                a += aHB[x * N3 + c] * iv * (-xf[x * N3 + c] + met[(M_XYZCC + x) * N3 + c]);
        }
        const int at = d == 0 ? i : d == 1 ? j : k;
        const int lo = d == 0 ? g.isg : d == 1 ? g.jsg : g.ksg;
        const bool oin = at - 1 >= lo && (d == 0 || i <= g.ieg) && (d == 1 || j <= g.jeg) && (d == 2 || k <= g.keg);
        if (oin) {
            const long cm = c - (d == 0 ? 1 : d == 1 ? (long)g.P : g.PP);
            const double iv = 1.0 / met[M_VOL * N3 + cm];
            for (int x = 0; x < 3; x++)
                a += aHB[x * N3 + cm] * iv * (xf[x * N3 + c] - met[(M_XYZCC + x) * N3 + cm]);
        }
        aS[(S_MF + d) * N3 + c] += (1.0 + w) * a;
        aO[(S_MF + d) * N3 + c] = -w * a;
        if (pw) pw[c] += (S[(S_MF + d) * N3 + c] - O[(S_MF + d) * N3 + c]) * a;
    }
}

// predictor gas transpose with the half-state gas adjoint (the EMF's
// corner density) included
__global__ __launch_bounds__(256) void k_predict_adj_mhd(Geo g, MParm m, Box cb, const Clock *clk,
                                                         const double *__restrict__ S,
                                                         const double *__restrict__ O,
                                                         const double *__restrict__ met,
                                                         const double *__restrict__ aHG,
                                                         const double *__restrict__ aWH,
                                                         const double *__restrict__ aWN,
                                                         double *__restrict__ aS,
                                                         double *__restrict__ aO,
                                                         double *__restrict__ pw)
{
    POINT_IN(cb);
    const long N3 = g.N3;
    const long c = IX(g, i, j, k);
    const double w = ab2_weight(clk);
    double Q[5], oQ[5], ahg[5], awh[5], awn[5], aQ[5] = {0, 0, 0, 0, 0}, aoQ[5] = {0, 0, 0, 0, 0}, aw = 0.0;
    for (int v = 0; v < 5; v++) {
        Q[v] = S[(S_GAS + v) * N3 + c];
        oQ[v] = O[(S_GAS + v) * N3 + c];
        ahg[v] = aHG[v * N3 + c];
        awh[v] = aWH[v * N3 + c];
        awn[v] = aWN[v * N3 + c];
    }
    predict_cell_adj(m, Q, oQ, w, met[M_VOL * N3 + c], ahg, awh, awn, aQ, aoQ, pw ? &aw : nullptr);
    if (pw) pw[c] = aw;
    for (int v = 0; v < 5; v++) {
        aS[v * N3 + c] += aQ[v];
        aO[v * N3 + c] = aoQ[v];
    }
    // oState's cell B does not enter the step (the predictor forms the half
    // state's B from the extrapolated face fluxes)
    for (int x = 0; x < 3; x++) aO[(S_B + x) * N3 + c] = 0.0;
}

// ------------------------------------------------------------ launchers
void launch_adj_objective(hipStream_t s, const Geo &g, const MParm &m, const double *S,
                          const double *wgt, double *aS, double *part, int *nparts)
{
    const Box cb = {g.isg, g.ieg, g.jsg, g.jeg, g.ksg, g.keg};
    const dim3 gr = agrid2(cb);
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
                       double *aSo, double *ag, const double *met, const double *gflx, double *pdt)
{
    const Box fb = {g.isg, g.ieg + 1, g.jsg, g.jeg + 1, g.ksg, g.keg + 1};
    hipLaunchKernelGGL(k_update_adj, agrid(fb), dim3(ABT), 0, s, g, fb, ab, clk, aSn, aSo, ag, met, gflx, pdt);
}

template <int D>
static void flux_adj_d(hipStream_t s, const Geo &g, const MParm &m, const Box &fb, const double *H,
                       const double *met, const double *ag, double *aWH, double *aWN)
{
    if (m.rmeth) hipLaunchKernelGGL((k_flux_adj<D, 1>), agrid(fb), dim3(ABT), 0, s, g, m, fb, H, met, ag, aWH, aWN);
    else hipLaunchKernelGGL((k_flux_adj<D, 0>), agrid(fb), dim3(ABT), 0, s, g, m, fb, H, met, ag, aWH, aWN);
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
                        double *aS, double *aO, double *pw)
{
    const Box cb = {g.isg, g.ieg, g.jsg, g.jeg, g.ksg, g.keg};
    hipLaunchKernelGGL(k_predict_adj, agrid(cb), dim3(ABT), 0, s, g, m, cb, clk, S, O, met, aWH, aWN,
                       aS, aO, pw);
}

void launch_set_dt(hipStream_t s, Clock *clk, const Clock *rec)
{
    hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(64), 0, s, clk, rec);
}

void launch_set_dt_val(hipStream_t s, Clock *clk, const double *dts, int n)
{
    hipLaunchKernelGGL(k_set_dt_val, dim3(1), dim3(64), 0, s, clk, dts, n);
}

static inline Box face_points(const Geo &g) { return Box{g.isg, g.ieg + 1, g.jsg, g.jeg + 1, g.ksg, g.keg + 1}; }

long dt_part_blocks(const Geo &g)
{
    const dim3 gr = agrid2(face_points(g));
    return (long)gr.x * gr.y * gr.z;
}

void launch_dt_fold(hipStream_t s, const Geo &g, const double *pdt, const double *pw, double *part,
                    const Clock *clk, int n, double *mu)
{
    const Box fb = face_points(g);
    const dim3 gr = agrid2(fb);
    hipLaunchKernelGGL(k_dt_part, gr, dim3(ABX, ABY), 0, s, g, fb, pdt, pw, part);
    hipLaunchKernelGGL(k_dt_fold, dim3(1), dim3(256), 0, s, part, (long)gr.x * gr.y * gr.z, clk, n, mu);
}

long dt_record_blocks(const Geo &g, const Box &dtb)
{
    const dim3 gr = agrid2(dtb);
    return (long)gr.x * gr.y * gr.z;
}

void launch_dt_record(hipStream_t s, const Geo &g, const MParm &m, const Box &dtb, const double *S,
                      const double *met, const double *bg, const Clock *clk, double *part, double *rec)
{
    const dim3 gr = agrid2(dtb);
    const long np = (long)gr.x * gr.y * gr.z;
    if (m.doBoris) {
        hipLaunchKernelGGL(k_dtmin_part<true>, gr, dim3(ABX, ABY), 0, s, g, m, dtb, S, met, bg, part);
        hipLaunchKernelGGL(k_dtmin_record<true>, dim3(1), dim3(256), 0, s, g, m, part, np, clk, S, met, bg, rec);
    } else {
        hipLaunchKernelGGL(k_dtmin_part<false>, gr, dim3(ABX, ABY), 0, s, g, m, dtb, S, met, bg, part);
        hipLaunchKernelGGL(k_dtmin_record<false>, dim3(1), dim3(256), 0, s, g, m, part, np, clk, S, met, bg, rec);
    }
}

void launch_dt_inject(hipStream_t s, const Geo &g, const MParm &m, const double *rec, const double *mu,
                      double *lam)
{
    if (m.doBoris) hipLaunchKernelGGL(k_dt_inject<true>, dim3(1), dim3(64), 0, s, g, m, rec, mu, lam);
    else hipLaunchKernelGGL(k_dt_inject<false>, dim3(1), dim3(64), 0, s, g, m, rec, mu, lam);
}

// ------------------------------------------------------------ MHD launchers
static inline Box cells_of(const Geo &g) { return Box{g.isg, g.ieg, g.jsg, g.jeg, g.ksg, g.keg}; }
static inline Box faces_of(const Geo &g) { return Box{g.isg, g.ieg + 1, g.jsg, g.jeg + 1, g.ksg, g.keg + 1}; }

void launch_cellb_adj(hipStream_t s, const Geo &g, int all, double *aS, const double *met)
{
    const Box cb = cells_of(g);
    hipLaunchKernelGGL(k_cellb_adj, agrid(cb), dim3(ABT), 0, s, g, cb, all, aS, met);
}

template <int SIDE>
static void bcp_adj_side(hipStream_t s, const Geo &g, double *aS)
{
    Box t;
    if (SIDE <= 1) { t.i0 = g.jsg; t.i1 = g.jeg + 1; t.j0 = g.ksg; t.j1 = g.keg + 1; }
    else if (SIDE <= 3) { t.i0 = g.isg; t.i1 = g.ieg + 1; t.j0 = g.ksg; t.j1 = g.keg + 1; }
    else { t.i0 = g.isg; t.i1 = g.ieg + 1; t.j0 = g.jsg; t.j1 = g.jeg + 1; }
    t.k0 = t.k1 = 0;
    const dim3 gr((unsigned)((t.i1 - t.i0 + 64) / 64), (unsigned)(t.j1 - t.j0 + 1), 9);
    hipLaunchKernelGGL(k_bc_periodic_adj<SIDE>, gr, dim3(64), 0, s, g, t, aS);
}

template <int SIDE, int KIND>
static void bcc_adj_side(hipStream_t s, const Geo &g, double *aS, const double *met)
{
    Box t;
    if (SIDE <= 1) { t.i0 = g.jsg; t.i1 = g.jeg; t.j0 = g.ksg; t.j1 = g.keg; }
    else if (SIDE <= 3) { t.i0 = g.isg; t.i1 = g.ieg; t.j0 = g.ksg; t.j1 = g.keg; }
    else { t.i0 = g.isg; t.i1 = g.ieg; t.j0 = g.jsg; t.j1 = g.jeg; }
    t.k0 = t.k1 = 0;
    const bool movesB = (SIDE == 0 && KIND == 2) || (SIDE == 1 && KIND == 3);
    const dim3 gr((unsigned)((t.i1 - t.i0 + 64) / 64), (unsigned)(t.j1 - t.j0 + 1), movesB ? 9 : 6);
    hipLaunchKernelGGL((k_bc_copy_adj<SIDE, KIND>), gr, dim3(64), 0, s, g, t, aS, met);
}

// transpose of one side's BC with magnetic fields; kind as gam_model.bc
void launch_bc_mhd_adj(hipStream_t s, int side, int kind, const Geo &g, double *aS, const double *met)
{
    if (kind == 0) { launch_bc_periodic_adj(s, side, g, aS); return; }
    switch (side) {
    case 0: if (kind == 2) bcc_adj_side<0, 2>(s, g, aS, met); else bcc_adj_side<0, 1>(s, g, aS, met); break;
    case 1: if (kind == 3) bcc_adj_side<1, 3>(s, g, aS, met); else bcc_adj_side<1, 1>(s, g, aS, met); break;
    case 2: bcc_adj_side<2, 1>(s, g, aS, met); break;
    case 3: bcc_adj_side<3, 1>(s, g, aS, met); break;
    case 4: bcc_adj_side<4, 1>(s, g, aS, met); break;
    default: bcc_adj_side<5, 1>(s, g, aS, met); break;
    }
}

void launch_bc_periodic_adj(hipStream_t s, int side, const Geo &g, double *aS)
{
    switch (side) {
    case 0: bcp_adj_side<0>(s, g, aS); break;
    case 1: bcp_adj_side<1>(s, g, aS); break;
    case 2: bcp_adj_side<2>(s, g, aS); break;
    case 3: bcp_adj_side<3>(s, g, aS); break;
    case 4: bcp_adj_side<4>(s, g, aS); break;
    default: bcp_adj_side<5>(s, g, aS); break;
    }
}

void launch_update_adj_mhd(hipStream_t s, const Geo &g, const MParm &m, const Box &ab, const Box &ctb,
                           const Box *eboxes, const Clock *clk, const double *So, const double *Sn1,
                           const double *gflx, const double *mflx, const double *met, const double *bg,
                           double *aSn, double *aSo, double *aUr, double *adM, double *ag, double *amf,
                           double *aE, double *pdt)
{
    const Box cb = cells_of(g), fb = faces_of(g);
    if (m.doBoris)
        hipLaunchKernelGGL(k_update_adj_cells<true>, agrid(cb), dim3(ABT), 0, s, g, m, cb, ab, clk, So, Sn1,
                           gflx, mflx, met, bg, m.doBG, aSn, aSo, aUr, adM, pdt);
    else
        hipLaunchKernelGGL(k_update_adj_cells<false>, agrid(cb), dim3(ABT), 0, s, g, m, cb, ab, clk, So, Sn1,
                           gflx, mflx, met, bg, m.doBG, aSn, aSo, aUr, adM, pdt);
    // the update's new cell B from its new face fluxes (every cell)
    launch_cellb_adj(s, g, 1, aSn, met);
    EdgeBoxes eb;
    for (int d = 0; d < 3; d++) eb.e[d] = eboxes[d];
    hipLaunchKernelGGL(k_update_adj_faces, agrid(fb), dim3(ABT), 0, s, g, fb, ab, ctb, eb, clk, met, aUr,
                       adM, aSn, aSo, ag, amf, aE);
}

template <int D, int RM>
static void flux_adj_mhd_r(hipStream_t s, const Geo &g, const MParm &m, const Box &fb, const double *H,
                           const double *S, const double *met, const double *bg, const double *ag,
                           const double *amf, double *sc, double *aHMF, double *aHG)
{
#define FLUXADJ(BG, BO)                                                                                      \
    hipLaunchKernelGGL((k_flux_adj_mhd<D, RM, BG, BO>), agrid_slab(fb), dim3(ABT), 0, s, g, m, fb, H, S, met,      \
                       bg, ag, amf, sc, aHMF, aHG)
    if (m.doBG && m.doBoris) FLUXADJ(true, true);
    else if (m.doBG) FLUXADJ(true, false);
    else if (m.doBoris) FLUXADJ(false, true);
    else FLUXADJ(false, false);
#undef FLUXADJ
}

template <int D>
static void flux_adj_mhd_d(hipStream_t s, const Geo &g, const MParm &m, const Box &fb, const double *H,
                           const double *S, const double *met, const double *bg, const double *ag,
                           const double *amf, double *sc, double *aHMF, double *aHG)
{
    if (m.rmeth) flux_adj_mhd_r<D, 1>(s, g, m, fb, H, S, met, bg, ag, amf, sc, aHMF, aHG);
    else flux_adj_mhd_r<D, 0>(s, g, m, fb, H, S, met, bg, ag, amf, sc, aHMF, aHG);
}

void launch_flux_adj_mhd(hipStream_t s, int d, const Geo &g, const MParm &m, const Box &fb, const double *H,
                         const double *S, const double *met, const double *bg, const double *ag,
                         const double *amf, double *sc, double *aHMF, double *aHG)
{
    if (d == 0) flux_adj_mhd_d<0>(s, g, m, fb, H, S, met, bg, ag, amf, sc, aHMF, aHG);
    else if (d == 1) flux_adj_mhd_d<1>(s, g, m, fb, H, S, met, bg, ag, amf, sc, aHMF, aHG);
    else flux_adj_mhd_d<2>(s, g, m, fb, H, S, met, bg, ag, amf, sc, aHMF, aHG);
}

void launch_recon_adj_gather(hipStream_t s, const Geo &g, const MParm &m, const double *sc, const double *rv,
                             const double *aVf, const double *met, double *aWH, double *aWN, double *aHB,
                             double *aSSB)
{
    const Box cb = cells_of(g);
    if (m.rmeth)
        hipLaunchKernelGGL(k_recon_adj_gather<1>, agrid_slab(cb), dim3(ABT), 0, s, g, cb, sc, rv, aVf, met,
                           m.doLFM ? 1 : 0, aWH, aWN, aHB, aSSB);
    else
        hipLaunchKernelGGL(k_recon_adj_gather<0>, agrid_slab(cb), dim3(ABT), 0, s, g, cb, sc, rv, aVf, met,
                           m.doLFM ? 1 : 0, aWH, aWN, aHB, aSSB);
}

template <int ED, int RM, bool BG, bool BO>
static void emf_adj_t(hipStream_t s, const Geo &g, const MParm &m, const Box &eb, const Clock *clk,
                      const double *H, const double *S, const double *Vf, const double *met, const double *bg,
                      const double *aE, double *sc, double *pdt)
{
    hipLaunchKernelGGL((k_emf_adj<ED, RM, BG, BO>), agrid_slab(eb), dim3(ABT), 0, s, g, m, eb, clk, H, S, Vf, met,
                       bg, aE, sc + (size_t)16 * ED * g.N3, pdt);
}

template <int ED>
static void emf_adj_d(hipStream_t s, const Geo &g, const MParm &m, const Box &eb, const Clock *clk,
                      const double *H, const double *S, const double *Vf, const double *met, const double *bg,
                      const double *aE, double *sc, double *pdt)
{
    const int sel = (m.rmeth ? 4 : 0) + (m.doBG ? 2 : 0) + (m.doBoris ? 1 : 0);
#define EMFADJ(RM, BG, BO) emf_adj_t<ED, RM, BG, BO>(s, g, m, eb, clk, H, S, Vf, met, bg, aE, sc, pdt)
    switch (sel) {
    case 0: EMFADJ(0, false, false); break;
    case 1: EMFADJ(0, false, true); break;
    case 2: EMFADJ(0, true, false); break;
    case 3: EMFADJ(0, true, true); break;
    case 4: EMFADJ(1, false, false); break;
    case 5: EMFADJ(1, false, true); break;
    case 6: EMFADJ(1, true, false); break;
    default: EMFADJ(1, true, true); break;
    }
#undef EMFADJ
}

void launch_emf_adj(hipStream_t s, int ed, const Geo &g, const MParm &m, const Box &eb, const Clock *clk,
                    const double *H, const double *S, const double *Vf, const double *met, const double *bg,
                    const double *aE, double *sc, double *pdt)
{
    if (ed == 0) emf_adj_d<0>(s, g, m, eb, clk, H, S, Vf, met, bg, aE, sc, pdt);
    else if (ed == 1) emf_adj_d<1>(s, g, m, eb, clk, H, S, Vf, met, bg, aE, sc, pdt);
    else emf_adj_d<2>(s, g, m, eb, clk, H, S, Vf, met, bg, aE, sc, pdt);
}

void launch_emf_adj_gather(hipStream_t s, const Geo &g, const MParm &m, const Box *eboxes, const double *sc,
                           const double *met, double *aVf, double *aHMF, double *aHG, double *aSSMF)
{
    const Box fb = faces_of(g);
    EdgeBoxes eb;
    for (int d = 0; d < 3; d++) eb.e[d] = eboxes[d];
    if (m.rmeth)
        hipLaunchKernelGGL(k_emf_adj_gather<1>, agrid_slab(fb), dim3(ABT), 0, s, g, fb, eb, sc, met,
                           m.doLFM ? 1 : 0, aVf, aHMF, aHG, aSSMF);
    else
        hipLaunchKernelGGL(k_emf_adj_gather<0>, agrid_slab(fb), dim3(ABT), 0, s, g, fb, eb, sc, met,
                           m.doLFM ? 1 : 0, aVf, aHMF, aHG, aSSMF);
}

void launch_vface_rv(hipStream_t s, const Geo &g, const double *met, double *rv)
{
    const Box cb = cells_of(g);
    hipLaunchKernelGGL(k_vface_rv, agrid(cb), dim3(ABT), 0, s, g, cb, met, rv);
}

void launch_predict_adj_mhd(hipStream_t s, const Geo &g, const MParm &m, const Clock *clk, const double *S,
                            const double *O, const double *met, const double *aHG, const double *aWH,
                            const double *aWN, double *aHB, double *aHMF, double *aS, double *aO, double *pw)
{
    const Box cb = cells_of(g), fb = faces_of(g);
    hipLaunchKernelGGL(k_predict_adj_mhd, agrid(cb), dim3(ABT), 0, s, g, m, cb, clk, S, O, met, aHG, aWH,
                       aWN, aS, aO, pw);
    hipLaunchKernelGGL(k_predict_mf_adj, agrid(fb), dim3(ABT), 0, s, g, fb, clk, met, aHB, aHMF, aS, aO, S, O,
                       pw);
}

// ============================================================ inflow BC
#include "gam_wind.h"

// Per-lane accumulators of the inflow-series adjoint: up to two sample
// indices with seven columns each (rho, vx, vy, vz, P, By, Bz).
struct WindAcc {
    int ia, ib;
    double a[7], b[7];
};

__device__ __forceinline__ void wacc_store(const WindDev &w, double *gQ, double *gB, int i0, const double *v)
{
    for (int q = 0; q < 5; q++)
        if (v[q] != 0.0) AADD(gQ + i0 + (long)q * w.n, v[q]);
    if (v[5] != 0.0) AADD(gB + i0 + (long)w.n, v[5]);
    if (v[6] != 0.0) AADD(gB + i0 + 2L * w.n, v[6]);
}

__device__ __forceinline__ void wacc_add(WindAcc &A, const WindDev &w, double *gQ, double *gB, int i0,
                                         const double *v)
{
    if (A.ia == i0 || A.ia < 0) {
        A.ia = i0;
        for (int q = 0; q < 7; q++) A.a[q] += v[q];
    } else if (A.ib == i0 || A.ib < 0) {
        A.ib = i0;
        for (int q = 0; q < 7; q++) A.b[q] += v[q];
    } else {
        wacc_store(w, gQ, gB, A.ia, A.a);
        A.ia = i0;
        for (int q = 0; q < 7; q++) A.a[q] = v[q];
    }
}

// Flush one accumulator slot of every lane of the (64-lane) block: the
// lanes holding the same sample are summed across the wave and one lane
// adds the sum, one set of atomics per distinct sample.  Every lane of the
// wave must call it.
__device__ __forceinline__ void wacc_flush(const WindDev &w, double *gQ, double *gB, int key, double *v)
{
    const int lane = (int)threadIdx.x;
    bool pend = key >= 0;
    for (;;) {
        const unsigned long long mk = __ballot(pend);
        if (mk == 0ull) break;
        const int lead = __builtin_ctzll(mk);
        const int k0 = __shfl(key, lead, 64);
        const bool mine = pend && key == k0;
        double sv[7];
        for (int q = 0; q < 7; q++) {
            double x = mine ? v[q] : 0.0;
            for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
            sv[q] = x;
        }
        if (lane == lead) wacc_store(w, gQ, gB, k0, sv);
        if (mine) pend = false;
    }
}

// Transpose of the inflow BC (k_bc_wind, wind.F90:226-335).  Ghost gas =
// P2C(wSW * wind sample + (1 - wSW) * last active cell's primitives, the
// latter with the outflow velocity raised to the fast speed); ghost face
// flux = wSW * A (n . B_wind) + (1 - wSW) * the first inner face's flux.
// Adds the adjoints of the last active cell (gas, and its cell B through
// the cell's face fluxes), of the inner faces, and of the inflow series
// samples (gQ (n,5), gB (n,3), column-major; the sample index is a step
// function of time and carries no derivative) into copy b % WIND_COPIES of
// the series adjoint for block b, which spreads the atomics of the many
// points that read one sample.  One thread per transverse point (j, k) and
// ghost layer; a 64-lane block holds 16 points along j with their four
// layers in adjacent lanes, so that a load reads 32 contiguous bytes (the
// layers along i) instead of 8.
__global__ __launch_bounds__(64) void k_bc_wind_adj(Geo g, MParm m, Box t, const Clock *clk, WindDev w,
                                                    int mhd, const double *__restrict__ S,
                                                    double *__restrict__ aS, const double *__restrict__ met,
                                                    double *__restrict__ gQ, double *__restrict__ gB)
{
    const int n = 1 + (int)(threadIdx.x & 3);   // ghost layer
    const int j = t.i0 + (int)(blockIdx.x * 16 + (threadIdx.x >> 2));
    const int k = t.j0 + (int)blockIdx.y;
    {
        const long cpy = (long)((blockIdx.x + blockIdx.y * gridDim.x) % WIND_COPIES);
        gQ += cpy * 5 * w.n;
        gB += cpy * 3 * w.n;
    }
    WindAcc A;
    A.ia = A.ib = -1;
    for (int q = 0; q < 7; q++) A.a[q] = A.b[q] = 0.0;
    const long N3 = g.N3;
    double *amf = aS + S_MF * N3;
    const int ip = g.ie;
    const long cp = IX(g, ip, j, k);
    const bool live = j <= t.i1 && k <= t.j1;
    const bool cell = live && k <= g.keg && j <= g.jeg;
    double aUp[5] = {0, 0, 0, 0, 0}, aBp[3] = {0, 0, 0};
    if (live) {
        const double *mf = S + S_MF * N3;
        const double tm = clk->t;
        double vF[3];
        wind_front(w, tm, vF);
        double U[5], Bi[3];
        for (int v = 0; v < 5; v++) U[v] = cell ? S[(S_GAS + v) * N3 + cp] : 0.0;
        for (int x = 0; x < 3; x++) Bi[x] = cell ? S[(S_B + x) * N3 + cp] : 0.0;
        {
            const int ig = g.ie + n;
            // ghost face fluxes
            const int fi[3] = {ig + 1, ig, ig};
            for (int d = 2; mhd && d >= 0; d--) {
                const long f = IX(g, fi[d], j, k);
                const double ao = amf[d * N3 + f];
                amf[d * N3 + f] = 0.0;
                if (ao == 0.0) continue;
                const double xf[3] = {met[(M_XFC + 0 + 3 * d) * N3 + f], met[(M_XFC + 1 + 3 * d) * N3 + f],
                                      met[(M_XFC + 2 + 3 * d) * N3 + f]};
                const double tl = wind_lag_time(w, vF, xf, tm);
                const int i0 = wind_index(w, tl);
                double r, p, V[3], B[3];
                wind_at(w, vF, xf, tm, &r, &p, V, B);
                const double wSW = wind_wgt_v(xf, V);
                const double *nrm = met + (long)(M_TF + 9 * d) * N3 + f;
                const double Af = met[(M_FACE + d) * N3 + f];
                const double sw = Af * (nrm[0] * B[0] + nrm[N3] * B[1] + nrm[2 * N3] * B[2]);
                const long fin = (d == 0) ? IX(g, ip + 1, j, k) : cp;
                const double in = mf[d * N3 + fin];
                AADD(amf + d * N3 + fin, (1.0 - wSW) * ao);
                double v7[7] = {0, 0, 0, 0, 0, 0, 0};
                if (tl > w.tMin) {
                    const double aBn = wSW * ao * Af;
                    const double aB[3] = {aBn * nrm[0], aBn * nrm[N3], aBn * nrm[2 * N3]};
                    v7[5] = aB[0] * w.ByC + aB[1];
                    v7[6] = aB[0] * w.BzC + aB[2];
                }
                double dw[3];
                wind_wgt_dv(w, vF, xf, tm, dw);
                const double awt = (sw - in) * ao;
                for (int x = 0; x < 3; x++) v7[1 + x] += awt * dw[x];
                wacc_add(A, w, gQ, gB, i0, v7);
            }
            // ghost gas
            double aU[5] = {0, 0, 0, 0, 0};
            bool any = false;
            const long cg = IX(g, ig, j, k);
            if (cell)
                for (int v = 0; v < 5; v++) {
                    aU[v] = aS[(S_GAS + v) * N3 + cg];
                    aS[(S_GAS + v) * N3 + cg] = 0.0;
                    any |= aU[v] != 0.0;
                }
            if (any) {
                const double xc[3] = {met[(M_XYZCC + 0) * N3 + cg], met[(M_XYZCC + 1) * N3 + cg],
                                      met[(M_XYZCC + 2) * N3 + cg]};
                const int i0 = wind_index(w, wind_lag_time(w, vF, xc, tm));
                double rs, ps, Vs[3], Bs[3];
                wind_at(w, vF, xc, tm, &rs, &ps, Vs, Bs);
                const double wSW = wind_wgt_v(xc, Vs), wMHD = 1.0 - wSW;
                double gs[5] = {0, 0, 0, 0, 0}, gi[5] = {0, 0, 0, 0, 0};
                if (wSW > GAM_TINY) {
                    gs[0] = rs; gs[1] = Vs[0]; gs[2] = Vs[1]; gs[3] = Vs[2]; gs[4] = ps;
                }
                double nH[3] = {0, 0, 0};
                if (wMHD > GAM_TINY) {
                    const long fp = IX(g, ip + 1, j, k);
                    for (int x = 0; x < 3; x++) nH[x] = met[(M_TF + x) * N3 + fp];
                    wind_inner(m, U, Bi, nH, gi);
                }
                double gw[5], agw[5] = {0, 0, 0, 0, 0};
                for (int v = 0; v < 5; v++) gw[v] = wSW * gs[v] + wMHD * gi[v];
                p2c_cell_adj(m, gw, aU, agw);
                // the series: the sample (when used) and the blend weight
                double v7[7] = {0, 0, 0, 0, 0, 0, 0};
                if (wSW > GAM_TINY)
                    for (int v = 0; v < 5; v++) v7[v] = wSW * agw[v];
                double awt = 0.0;
                for (int v = 0; v < 5; v++) awt += (gs[v] - gi[v]) * agw[v];
                double dw[3];
                wind_wgt_dv(w, vF, xc, tm, dw);
                for (int x = 0; x < 3; x++) v7[1 + x] += awt * dw[x];
                wacc_add(A, w, gQ, gB, i0, v7);
                if (wMHD > GAM_TINY) {
                    // the interior cell
                    double agi[5];
                    for (int v = 0; v < 5; v++) agi[v] = wMHD * agw[v];
                    wind_inner_adj(m, U, Bi, nH, agi, aUp, aBp);
                }
            }
        }
    }
    // the four layers' lanes share the last active cell: summed across them,
    // added by the first (no other thread writes that cell's gas adjoint)
    for (int v = 0; v < 5; v++) {
        aUp[v] += __shfl_xor(aUp[v], 1, 64);
        aUp[v] += __shfl_xor(aUp[v], 2, 64);
    }
    for (int x = 0; x < 3; x++) {
        aBp[x] += __shfl_xor(aBp[x], 1, 64);
        aBp[x] += __shfl_xor(aBp[x], 2, 64);
    }
    if (cell && n == 1) {
        for (int v = 0; v < 5; v++)
            if (aUp[v] != 0.0) aS[(S_GAS + v) * N3 + cp] += aUp[v];
        if (mhd && (aBp[0] != 0.0 || aBp[1] != 0.0 || aBp[2] != 0.0)) {
            // the last active cell's B = its bFlux2Fld from its face fluxes
            // (faces shared with the neighbouring points: atomic adds)
            double af[6] = {0, 0, 0, 0, 0, 0};
            cell_b_adj(g, met, cp, aBp, af);
            AADD(amf + cp, af[0]);
            AADD(amf + N3 + cp, af[1]);
            AADD(amf + 2 * N3 + cp, af[2]);
            AADD(amf + cp + 1, af[3]);
            AADD(amf + N3 + cp + g.P, af[4]);
            AADD(amf + 2 * N3 + cp + g.PP, af[5]);
        }
    }
    wacc_flush(w, gQ, gB, A.ia, A.a);
    wacc_flush(w, gQ, gB, A.ib, A.b);
}

void launch_bc_wind_adj(hipStream_t s, const Geo &g, const MParm &m, const Clock *clk, const WindDev &w,
                        const double *S, double *aS, const double *met, double *gQ, double *gB)
{
    // transverse ranges of the forward (launch_bc, kind 4)
    Box t;
    t.i0 = g.jsg; t.i1 = g.jeg + 1; t.j0 = g.ksg; t.j1 = g.keg + 1; t.k0 = t.k1 = 0;
    const dim3 gr((unsigned)((t.i1 - t.i0 + 16) / 16), (unsigned)(t.j1 - t.j0 + 1), 1);
    hipLaunchKernelGGL(k_bc_wind_adj, gr, dim3(64), 0, s, g, m, t, clk, w, m.doMHD ? 1 : 0, S, aS, met, gQ, gB);
}
