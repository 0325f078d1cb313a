// gam_step.hip -- gfx950 kernels of one GAMERA MHD step (FP64).
//
// Kernel sequence of one stepGamera (gamapp.F90:73-113):
//   k_predict    Predictor: AB2 face fluxes, half-step B and gas, plus the
//                primitives used by the flux sweeps
//   k_flux       Lyon gas/magnetic fluxes, one per direction, and the face
//                velocities (Mom2Face) of the same faces
//   k_vface_rim  face velocities the EMF stencils read past the flux faces
//   k_emf        edge EMFs
//   k_update     CT face-flux update, cell B of the new state, flux
//                divergence, Reynolds, Maxwell/Boris and the CFL
//                min-reduction of the new state
//   k_finalize   time bookkeeping and CalcDT checks (1 thread)
//   k_bc_*       boundary conditions
//   k_bflux2fld  cell B after the BCs; active cells keep theirs,
//                since no BC writes a face of an active cell
// State and oState are two buffers that swap roles every step, which
// replaces CopyState (mhdgroup.F90:204-255): the update reads the current
// buffer and writes every point of the other.
#include "gam_phys.h"
#include "gam_kernels.h"
#include <type_traits>

#define BX 64
#define BY 4

__device__ __forceinline__ bool in_box(const Box &b, int i, int j, int k)
{
    return i >= b.i0 && i <= b.i1 && j >= b.j0 && j <= b.j1 && k >= b.k0 && k <= b.k1;
}

#define THREAD_IJK(b)                                   \
    const int i = (b).i0 + blockIdx.x * BX + threadIdx.x; \
    const int j = (b).j0 + blockIdx.y * BY + threadIdx.y; \
    const int k = (b).k0 + blockIdx.z;                  \
    if (i > (b).i1 || j > (b).j1) return;

// XCD-aware tiling.  A launch covers box b row by row (a row = the points
// of one (j, k) along i) and cuts the rows, i fastest, into chunks of 256
// consecutive points, one per workgroup, so no lane idles however the row
// length falls against 64.  Row order, by ax1:
//   1: j fastest, then k;   2: k fastest, then j;
//   3: slabs of RB j-rows: j fastest within the slab, then k, then the
//      next slab -- the rows at j+1 and at k+1 are both at most RB rows
//      away, for kernels that read neighbours along j and along k.
// The hardware hands workgroup w to XCD w % 8, so logical chunk
// L = (w % 8) * per + w / 8 gives every XCD one contiguous run of rows:
// chunks that share stencil rows run close together on the same L2.
#define RB 8
struct Tile {
    Box b;
    int ax1;
    int ni, nj, nk;   // extents
    int nb;           // chunks in all
    int ch;           // points per chunk (= threads per workgroup)
};

__device__ __forceinline__ bool tile_point(const Tile &t, int w, int &i, int &j, int &k)
{
    const int per = (t.nb + 7) >> 3;
    const int L = (w & 7) * per + (w >> 3);
    i = j = k = 0x40000000;
    if (L >= t.nb) return false;
    const unsigned p = (unsigned)L * (unsigned)t.ch + threadIdx.x + BX * threadIdx.y;
    if (p >= (unsigned)t.ni * (unsigned)(t.nj * t.nk)) return false;
    const unsigned R = p / (unsigned)t.ni;
    i = t.b.i0 + (int)(p - R * (unsigned)t.ni);
    int jj, kk;
    if (t.ax1 == 1) {
        kk = (int)(R / (unsigned)t.nj); jj = (int)(R - (unsigned)kk * t.nj);
    } else if (t.ax1 == 2) {
        jj = (int)(R / (unsigned)t.nk); kk = (int)(R - (unsigned)jj * t.nk);
    } else {
        const unsigned slab = (unsigned)RB * t.nk;
        const unsigned nfull = (unsigned)(t.nj / RB);
        const unsigned sb = R / slab;
        if (sb < nfull) {
            const unsigned r = R - sb * slab;
            kk = (int)(r / RB);
            jj = (int)(sb * RB + (r % RB));
        } else {
            const unsigned r = R - nfull * slab;
            const unsigned rj = (unsigned)t.nj - nfull * RB;
            kk = (int)(r / rj);
            jj = (int)(nfull * RB + (r - (unsigned)kk * rj));
        }
    }
    j = t.b.j0 + jj;
    k = t.b.k0 + kk;
    return true;
}

#define TILE_IJK(t)                  \
    int i, j, k;                     \
    if (!tile_point((t), (int)bi.x, i, j, k)) return;

static inline Tile tile_of(const Box &b, int ax1, int ch = BX * BY)
{
    Tile t;
    t.b = b;
    t.ax1 = ax1;
    t.ni = b.i1 - b.i0 + 1;
    t.nj = b.j1 - b.j0 + 1;
    t.nk = b.k1 - b.k0 + 1;
    t.ch = ch;
    const long pts = (long)t.ni * t.nj * t.nk;
    t.nb = (int)((pts + ch - 1) / ch);
    return t;
}

// A launch plan value of a chunk-tiled slot: the row order in its low 3
// bits, the chunk size (threads per workgroup) above them.
static inline int plan_ax(int v) { return v & 7; }
static inline int plan_chunk(int v) { return (v >> 3) == 1 ? 512 : (v >> 3) == 2 ? 128 : BX * BY; }
static inline dim3 tgrid(const Tile &t) { return dim3((unsigned)(((t.nb + 7) / 8) * 8)); }

static inline dim3 grid_of(const Box &b)
{
    return dim3((unsigned)((b.i1 - b.i0 + BX) / BX), (unsigned)((b.j1 - b.j0 + BY) / BY),
                (unsigned)(b.k1 - b.k0 + 1));
}

// ====================================================================
// Launch forms.  Every kernel of the step is a functor whose body takes its
// workgroup's block index bi.  k_one<F> runs F over its own grid (one
// context); k_ens<F> runs F for many contexts in one launch: the
// concatenated grid of the members, each member's workgroups starting at a
// multiple of 8 (so a member's chunk-to-XCD mapping is the same as in its
// own launch) and each with its own argument record.  A table after the
// header names the member of every run of 8 workgroups.
// ====================================================================
// Threads per workgroup of functor F: F::NT where it declares one, else 256.
template <class F, class = void> struct nthreads { static constexpr int v = BX * BY; };
template <class F> struct nthreads<F, std::void_t<decltype(F::NT)>> { static constexpr int v = F::NT; };

template <class F, class... A>
__global__ __launch_bounds__(nthreads<F>::v) void k_one(A... a)
{
    F{}(make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), a...);
}

// Pointers read from an argument record are generic pointers to the
// compiler; every one of them addresses device (global) memory, and saying
// so lets it use global instead of flat loads and stores.
template <class T>
__device__ __forceinline__ const T &in_global(const T &v)
{
    return v;
}
template <class T>
__device__ __forceinline__ T *in_global(T *const &p)
{
    __attribute__((address_space(1))) T *q = (__attribute__((address_space(1))) T *)p;
    asm("" : "+s"(q));   // keeps the cast from folding back to a generic pointer
    return (T *)q;
}

// Call F with the record's fields, in order (passed by reference into the
// record, so that arrays in them are read where they lie).
template <class F, class... B>
__device__ __forceinline__ void pcall(const Pack<> &, uint3 bi, const B &...b)
{
    F{}(bi, in_global(b)...);
}
template <class F, class A0, class... A, class... B>
__device__ __forceinline__ void pcall(const Pack<A0, A...> &p, uint3 bi, const B &...b)
{
    pcall<F>(p.r, bi, b..., p.a0);
}

template <class F, class... A>
__global__ __launch_bounds__(nthreads<F>::v) void k_ens(const EnsHdr *__restrict__ h,
                                             const Pack<A...> *__restrict__ rec)
{
    const unsigned w = blockIdx.x;
    // owner of each run of 8 workgroups, stored right after the header
    const unsigned lo = reinterpret_cast<const unsigned short *>(h + 1)[w >> 3];
    const unsigned gx = h->gx[lo], gy = h->gy[lo], gz = h->gz[lo];
    unsigned l = w - h->wg0[lo];
    if (l >= gx * gy * gz) return;   // padding of the member's range
    uint3 bi;
    bi.x = l % gx;
    l /= gx;
    bi.y = l % gy;
    bi.z = l / gy;
    pcall<F>(rec[lo], bi);
}

template <class F, class... A>
static void launch_ens(hipStream_t s, const void *hdr, const void *rec, unsigned nwg, dim3 block)
{
    hipLaunchKernelGGL((k_ens<F, A...>), dim3(nwg), block, 0, s, (const EnsHdr *)hdr,
                       (const Pack<A...> *)rec);
}

// Batchable launch: issued at once, or appended to this thread's recorder.
static thread_local std::vector<EnsLaunch> *g_rec = nullptr;
void set_launch_recorder(std::vector<EnsLaunch> *r) { g_rec = r; }

template <class F, class... A>
static void launch_f(dim3 grid, dim3 block, hipStream_t s, A... a)
{
    if (g_rec) {
        const Pack<A...> p = make_pack(a...);
        EnsLaunch e;
        e.launch = &launch_ens<F, A...>;
        e.grid = grid;
        e.block = block;
        e.rec.resize(sizeof(p));
        memcpy(e.rec.data(), &p, sizeof(p));
        e.stream = s;
        e.mark = 0;
        e.ev = -1;
        g_rec->push_back(std::move(e));
        return;
    }
    hipLaunchKernelGGL((k_one<F, A...>), grid, block, 0, s, a...);
}

static void order_mark(hipStream_t s, int mark, int ev)
{
    EnsLaunch e;
    e.launch = nullptr;
    e.grid = e.block = dim3(0, 0, 0);
    e.stream = s;
    e.mark = mark;
    e.ev = ev;
    g_rec->push_back(std::move(e));
}

void order_record(hipStream_t s, hipEvent_t e, int ev)
{
    if (g_rec) order_mark(s, 1, ev);
    else hipEventRecord(e, s);
}

void order_wait(hipStream_t s, hipEvent_t e, int ev)
{
    if (g_rec) order_mark(s, 2, ev);
    else hipStreamWaitEvent(s, e, 0);
}


// ====================================================================
// Predictor (mhdgroup.F90:444-561) fused with bFlux2Fld of the half state
// (gridutils.F90:281-301) and the primitive conversions of the flux
// stencils (BlockCon2Prim, recon.F90:75-90) for both the half state and
// the limiter state n.
// Domain: all face points isg..ieg+1 (x) jsg..jeg+1 (x) ksg..keg+1.
// ====================================================================
template <int NTH = BX * BY>
struct k_predict {
    static constexpr int NT = NTH;
    __device__ __forceinline__ void operator()(const uint3 bi, Geo g, MParm m,
                                               Tile fb, const Clock *clk,
                                               const double *__restrict__ S,
                                               const double *__restrict__ O,
                                               const double *__restrict__ met,
                                               double *__restrict__ H) const
    {
        TILE_IJK(fb);
        const long N3 = g.N3;
        const long c = IX(g, i, j, k);
        const unsigned oc = (unsigned)(c * 8);    // byte offset of the point
        const double st = clk->st, dtM = clk->dt;
        double ost = clk->ost;
        if (fabs(ost - st) < GAM_TINY) ost = st - dtM;   // mhdgroup.F90:26-28
        const double odt = st - ost;
        const double pdt = 0.5 * clk->dtA;
        const double w = pdt / odt;

        double f[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            f[d] = m.doMHD ? ab2(ldp(S + (S_MF + d) * N3, oc), ldp(O + (S_MF + d) * N3, oc), w)
                           : ldp(H + (H_MF + d) * N3, oc);
            stp(H + (H_MF + d) * N3, oc, f[d]);
        }
        if (i > g.ieg || j > g.jeg || k > g.keg) return;

        // half-state cell B from the six extrapolated face fluxes
        double fIp, fJp, fKp;
        const unsigned oI = oc + 8, oJ = oc + 8 * g.P, oK = oc + (unsigned)(8 * g.PP);
        if (m.doMHD) {
            fIp = ab2(ldp(S + (S_MF + 0) * N3, oI), ldp(O + (S_MF + 0) * N3, oI), w);
            fJp = ab2(ldp(S + (S_MF + 1) * N3, oJ), ldp(O + (S_MF + 1) * N3, oJ), w);
            fKp = ab2(ldp(S + (S_MF + 2) * N3, oK), ldp(O + (S_MF + 2) * N3, oK), w);
        } else {
            fIp = ldp(H + (H_MF + 0) * N3, oI);
            fJp = ldp(H + (H_MF + 1) * N3, oJ);
            fKp = ldp(H + (H_MF + 2) * N3, oK);
        }
        double B[3];
        cell_b(g, met, c, f[0], f[1], f[2], fIp, fJp, fKp, B);
#pragma unroll
        for (int x = 0; x < 3; x++) stp(H + (H_B + x) * N3, oc, B[x]);

        // PredictCell (mhdgroup.F90:528-561)
        double Q[5], oQ[5], W[5], oW[5], pW[5], pQ[5], Wb[5];
#pragma unroll
        for (int v = 0; v < 5; v++) { Q[v] = ldp(S + (S_GAS + v) * N3, oc); oQ[v] = ldp(O + (S_GAS + v) * N3, oc); }
        c2p_cell(m, oQ, oW);
        c2p_cell(m, Q, W);
#pragma unroll
        for (int v = 0; v < 5; v++) pW[v] = ab2(W[v], oW[v], w);
        pW[DEN] = dmax(pW[DEN], m.dFloor);
        pW[PRS] = dmax(pW[PRS], m.pFloor);
        p2c_cell(m, pW, pQ);
#pragma unroll
        for (int v = 0; v < 5; v++) stp(H + (H_GAS + v) * N3, oc, pQ[v]);
        // primitives of the reconstruction stencils (half state), stored
        // multiplied by the cell volume: every use is the volume-weighted
        // interpolation <V q>/<V> (recon.F90:125-198, fields.F90:262-270), and
        // the product is the one those formulas would form.  The limiter
        // stencils take plain primitives: of state n (doLFMLim), else of the
        // half state.
        c2p_block(m, pQ, Wb);
        const double vol = ldp(met + M_VOL * N3, oc);
#pragma unroll
        for (int v = 0; v < 5; v++) stp(H + (H_WH + v) * N3, oc, vol * Wb[v]);
        if (m.doLFM) c2p_block(m, Q, Wb);
#pragma unroll
        for (int v = 0; v < 5; v++) stp(H + (H_WN + v) * N3, oc, Wb[v]);
    }
};

// ====================================================================
// The rest of Mom2Face.  The interface-flux kernels write the face
// velocities of their own faces; the EMF stencils (GetCornerV, 8 faces
// along dT2) also reach 4 faces past them on each side:
//   I faces at j in js-4..js-1 and je+1..je+4 (k-edges, dT2 = J),
//   J faces at k in ks-4..ks-1 and ke+1..ke+4 (i-edges, dT2 = K),
//   K faces at i in is-4..is-1 and ie+1..ie+4 (j-edges, dT2 = I).
// One launch covers the six slabs; slab q is face direction q / 2.
// ====================================================================
struct RimSet {
    Tile t[6];
    int start[7];    // first workgroup of each slab
};

struct k_vface_rim {
    __device__ __forceinline__ void operator()(const uint3 bi, Geo g, RimSet r,
                                               const double *__restrict__ H,
                                               const double *__restrict__ met,
                                               double *__restrict__ Vf) const
    {
        const int w = (int)bi.x;
        int q = 0;
#pragma unroll
        for (int p = 1; p < 6; p++) q += (w >= r.start[p]) ? 1 : 0;
        int i, j, k;
        if (!tile_point(r.t[q], w - r.start[q], i, j, k)) return;
        const int d = q >> 1;
        const long N3 = g.N3;
        const long c = IX(g, i, j, k);
        const long s = d == 0 ? 1L : (d == 1 ? (long)g.P : g.PP);
        const double *vol = met + (long)M_VOL * N3;
        const double *wv = H + (long)(H_WH + 1) * N3;   // V * velocity
        double v8[8];
#pragma unroll
        for (int n = 0; n < 8; n++) v8[n] = vol[c + (n - 4) * s];
        const double Vi = wsum8([&](int n) { return v8[n]; });
        const double rVi = 1.0 / Vi;
#pragma unroll
        for (int x = 0; x < 3; x++) {
            double q8[8];
#pragma unroll
            for (int n = 0; n < 8; n++) q8[n] = wv[x * N3 + c + (n - 4) * s];
            Vf[(3 * d + x) * N3 + c] = mom2face(q8, Vi, rVi);
        }
    }
};

// ====================================================================
// Edge EMF: GetCornerV/GetCornerD/GetCornerB/EdgeLRs + VxB
// (fields.F90:187-224, 288-517).  One thread per active edge of ED.
// ====================================================================
template <int FD, int SD, int RM>
__device__ __forceinline__ void edge_lr(const Geo &g, const MParm &m, long c,
                                        const double *__restrict__ met,
                                        const double *__restrict__ mfH,
                                        const double *__restrict__ mfL,
                                        double &bL, double &bR, double &fA)
{
    // EdgeLRs (fields.F90:288-362) for FD faces stacked along SD, area weighted
    const long N3 = g.N3;
    const long s = STRIDE<SD>(g);
    double A[8], q[8], nq[4];
#pragma unroll
    for (int n = 0; n < 8; n++) {
        A[n] = ldx<SD>(met + (M_FACE + FD) * N3, c, s, n - 4);
        q[n] = ldx<SD>(mfH + FD * N3, c, s, n - 4) / dmax(A[n], GAM_TINY);
    }
    // the limiter faces are the middle four of the stencil
#pragma unroll
    for (int n = 0; n < 4; n++) nq[n] = ldx<SD>(mfL + FD * N3, c, s, n - 2) / dmax(A[n + 2], GAM_TINY);
    double Vi = wsum8([&](int n) { return A[n]; });
    double l, r;
    recon_lr<RM>(m.pdmb, A, 1.0 / Vi, nq, q, l, r);
    fA = Vi;
    bL = l * fA;
    bR = r * fA;
}

template <int ED, int RM, bool BG, bool BORIS, int NTH = BX * BY>
struct k_emf {
    static constexpr int NT = NTH;
    __device__ __forceinline__ void operator()(const uint3 bi, Geo g, MParm m,
                                               Tile b, const Clock *clk,
                                               const double *__restrict__ H,
                                               const double *__restrict__ S,
                                               const double *__restrict__ Vf,
                                               const double *__restrict__ met,
                                               const double *__restrict__ bg,
                                               double *__restrict__ E) const
    {
        TILE_IJK(b);
        constexpr int T1 = (ED == 0) ? 1 : (ED == 1) ? 2 : 0;
        constexpr int T2 = (ED == 0) ? 2 : (ED == 1) ? 0 : 1;
        // getNormalPlane (metric.F90:151-167)
        constexpr int E1 = (ED == 0) ? 1 : 0;
        constexpr int E2 = (ED == 2) ? 1 : 2;
        const long N3 = g.N3;
        const long c = IX(g, i, j, k);
        const long s2 = STRIDE<T2>(g);

        // GetCornerV (fields.F90:457-517), doVdA
        double A[8];
#pragma unroll
        for (int n = 0; n < 8; n++) A[n] = ldx<T2>(met + (M_FACE + T1) * N3, c, s2, n - 4);
        const double rdA = 1.0 / dmax(wsum8([&](int n) { return A[n]; }), GAM_TINY);
        double vc[3];
#pragma unroll
        for (int d = 0; d < 3; d++)
            vc[d] = wsum8([&](int n) { return A[n] * ldx<T2>(Vf + d * N3, c, s2, n - 4); }) * rdA;
        const double *Te = met + (long)(M_TE + 9 * ED) * N3 + c;
        const double v1 = vc[0] * Te[3 * N3] + vc[1] * Te[4 * N3] + vc[2] * Te[5 * N3];
        const double v2 = vc[0] * Te[6 * N3] + vc[1] * Te[7 * N3] + vc[2] * Te[8 * N3];
        double vDiff = sqrt(v1 * v1 + v2 * v2);

        // GetCornerD (fields.F90:427-454)
        const long se1 = STRIDE<E1>(g), se2 = STRIDE<E2>(g);
        const double *Dh = H + (H_GAS + DEN) * N3;
        const double Dc = 0.25 * (Dh[c] + Dh[c - se1] + Dh[c - se2] + Dh[c - se1 - se2]);

        // GetCornerB (fields.F90:365-424)
        const double *mfH = H + H_MF * N3;
        const double *mfL = m.doLFM ? S + S_MF * N3 : mfH;
        double bT1L, bT1R, fA1, bT2L, bT2R, fA2;
        edge_lr<T1, T2, RM>(g, m, c, met, mfH, mfL, bT1L, bT1R, fA1);
        edge_lr<T2, T1, RM>(g, m, c, met, mfH, mfL, bT2L, bT2R, fA2);
        const double *Tb = met + (long)(M_TEB + 4 * ED) * N3 + c;
        const double xni = Tb[0], yni = Tb[N3], xnj = Tb[2 * N3], ynj = Tb[3 * N3];
        const double detT = 1.0 / (xni * ynj - xnj * yni);
        const double rA1 = 1.0 / fA1, rA2 = 1.0 / fA2;
        const double bT1 = 0.5 * (bT1L + bT1R) * rA1;
        const double bT2 = 0.5 * (bT2L + bT2R) * rA2;
        const double db1 = (bT1R - bT1L) * rA1;
        const double db2 = (bT2R - bT2L) * rA2;
        const double Jd = db2 - db1;
        double b1 = detT * (ynj * bT1 - yni * bT2);
        double b2 = detT * (-(xnj * bT1) + xni * bT2);
        if (BG) {
            b1 = b1 + bg[(B_EDGB0 + 0 + 2 * ED) * N3 + c];
            b2 = b2 + bg[(B_EDGB0 + 1 + 2 * ED) * N3 + c];
        }
        const double edge = met[(M_EDGE + ED) * N3 + c];
        double vA = sqrt(b1 * b1 + b2 * b2) / sqrt(Dc);
        if (BORIS) vA = m.Ca * vA / sqrt(m.Ca * m.Ca + vA * vA);
        vDiff = vDiff + vA;
        vDiff = dmin(vDiff, m.CFL * edge / clk->dt);
        const double e = -(v1 * b2 - v2 * b1) + m.Vd0 * vDiff * Jd;
        E[ED * N3 + c] = e * edge;
    }
};

// ====================================================================
// The fluxes of one face from its reconstructed L/R states (the part of
// Fluxes, stress.F90:204-429, after BlockStateLRs): magnetic kinetic flux
// with the B0 terms, Boris mag-HOGS, gas kinetic flux, HOGS and the nuke
// term, times the face area, stored to gFlx/mFlx of direction D.  AT(p)
// reads plane p at the face point, AT_M(p) one cell below it along D;
// RECON_B(BL, BR) fills the magnetic L/R states (called only with MHD on,
// after the metric loads, which keeps the register peak of the fused
// kernel at that of the gas part); oc is the face's byte offset for the
// stores.
// ====================================================================
template <int D, bool BG, bool BORIS, class AtF, class AtMF, class ReconB>
__device__ __forceinline__ void face_flux_lr(const Geo &g, const MParm &m,
                                             const double *__restrict__ H,
                                             const double *__restrict__ met,
                                             const double *__restrict__ bg, AtF AT, AtMF AT_M,
                                             ReconB RECON_B, const double *WL,
                                             const double *WR, double *BL, double *BR,
                                             unsigned oc,
                                             double *__restrict__ gflx,
                                             double *__restrict__ mflx)
{
    const long N3 = g.N3;
    double T[9];
#pragma unroll
    for (int q = 0; q < 9; q++) T[q] = AT(met + (M_TF + q + 9 * D) * N3);
    const double faB = AT(met + (M_FACE + D) * N3);

    double VaD = 0.0;
    if (m.doMHD) {
        RECON_B(BL, BR);
        const double iFaB = 1.0 / faB;
        const double Bn = AT(H + (H_MF + D) * N3) * iFaB;
        double B0[3] = {0.0, 0.0, 0.0}, B0n = 0.0;
        if (BG) {
#pragma unroll
            for (int x = 0; x < 3; x++) B0[x] = AT(bg + (B_FCB0 + x + 3 * D) * N3);
            B0n = AT(bg + (B_BFLUX0 + D) * N3) * iFaB;
        }
        double mF[3] = {0.0, 0.0, 0.0}, bbD = 0.0;
        mag_kin<BG, BORIS>(m, -1.0, WL, BL, Bn, T, B0, B0n, mF, bbD, VaD);
        mag_kin<BG, BORIS>(m, +1.0, WR, BR, Bn, T, B0, B0n, mF, bbD, VaD);
        if (BORIS) {
#pragma unroll
            for (int x = 0; x < 3; x++) {
                double deeP = bbD * (WR[1 + x] - WL[1 + x]) * m.iCa2;
                mF[x] = mF[x] - m.cHogM * VaD * deeP;
            }
        }
#pragma unroll
        for (int x = 0; x < 3; x++) stp(mflx + (x + 3 * D) * N3, oc, faB * mF[x]);
    }

    double gF[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, dW[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    gas_kin(m.igm1, -1.0, WL, T, gF, dW);
    gas_kin(m.igm1, +1.0, WR, T, gF, dW);
#pragma unroll
    for (int v = 0; v < 5; v++) gF[v] = gF[v] - m.cHogH * VaD * dW[v];
    if (BORIS) {   // "nuclear option" (stress.F90:399-417)
        double vL = sqrt(WL[1] * WL[1] + WL[2] * WL[2] + WL[3] * WL[3]);
        double vR = sqrt(WR[1] * WR[1] + WR[2] * WR[2] + WR[3] * WR[3]);
        double vn = sqrt(vL * vL + vR * vR) / GAM_SQRT2_SP;
        if (vn >= 1.5 * m.Ca) {
#pragma unroll
            for (int v = 0; v < 5; v++) {
                double dU = AT(H + (H_GAS + v) * N3) - AT_M(H + (H_GAS + v) * N3);
                gF[v] = gF[v] - m.Ca * dU;
            }
        }
    }
#pragma unroll
    for (int v = 0; v < 5; v++) stp(gflx + (v + 5 * D) * N3, oc, faB * gF[v]);
}

// ====================================================================
// Interface fluxes of one direction (Fluxes, stress.F90:204-429), every
// stencil point read from global memory (used along i, where the 8 points
// are contiguous and merge into wide loads)
// ====================================================================
template <int D, int RM, bool BG, bool BORIS, int NTH = BX * BY>
struct k_flux {
    static constexpr int NT = NTH;
    __device__ __forceinline__ void operator()(const uint3 bi, Geo g, MParm m,
                                               Tile b, const double *__restrict__ H,
                                               const double *__restrict__ S,
                                               const double *__restrict__ met,
                                               const double *__restrict__ bg,
                                               double *__restrict__ gflx,
                                               double *__restrict__ mflx,
                                               double *__restrict__ Vf) const
    {
        TILE_IJK(b);
        const long N3 = g.N3;
        const long c = IX(g, i, j, k);
        const long s = STRIDE<D>(g);
        // byte offsets of the 8 stencil points (cells c-4s .. c+3s), shared by
        // every plane; o[4] is the face point c itself.  Along i the points are
        // contiguous and the compiler merges them into wide loads from one
        // address, so that direction keeps plain indexing.
        unsigned o[8];
#pragma unroll
        for (int n = 0; n < 8; n++) o[n] = (unsigned)((c + (n - 4) * s) * 8);
        auto LD = [&](const double *__restrict__ plane, int n) -> double {
            if (D == 0) return plane[c + (n - 4)];
            return ldp(plane, o[n]);
        };

        double vol[8];
#pragma unroll
        for (int n = 0; n < 8; n++) vol[n] = LD(met + M_VOL * N3, n);
        const double Vi = wsum8([&](int n) { return vol[n]; });
        const double rVi = 1.0 / Vi;

        // L/R primitives (BlockStateLRs, recon.F90:39-93); Wv holds V*W
        const double *Wv = H + H_WH * N3;
        const double *Wl = H + H_WN * N3;
        double WL[5], WR[5];
#pragma unroll
        for (int v = 0; v < 5; v++) {
            double qv[8], nq[4];
#pragma unroll
            for (int n = 0; n < 8; n++) qv[n] = LD(Wv + v * N3, n);
#pragma unroll
            for (int n = 0; n < 4; n++) nq[n] = LD(Wl + v * N3, n + 2);
            recon_lr_w<RM>(m.pdmb, rVi, nq, qv, WL[v], WR[v]);
            // Mom2Face of this face from the same velocity stencil
            // (written without a condition on Vf: a branch here keeps the
            // stencils alive and costs ~30 VGPRs)
            if (v >= 1 && v <= 3) stp(Vf + (v - 1) * N3, o[4], mom2face(qv, Vi, rVi));
        }
        auto recon_b = [&](double *BL, double *BR) {
            const double *Bh = H + H_B * N3;
            const double *Bl = m.doLFM ? S + S_B * N3 : Bh;
#pragma unroll
            for (int x = 0; x < 3; x++) {
                double q[8], nq[4];
#pragma unroll
                for (int n = 0; n < 8; n++) q[n] = LD(Bh + x * N3, n);
#pragma unroll
                for (int n = 0; n < 4; n++) nq[n] = LD(Bl + x * N3, n + 2);
                recon_lr<RM>(m.pdmb, vol, rVi, nq, q, BL[x], BR[x]);
            }
        };
        double BL[3], BR[3];
        face_flux_lr<D, BG, BORIS>(
            g, m, H, met, bg, [&](const double *p) { return LD(p, 4); },
            [&](const double *p) { return LD(p, 3); }, recon_b, WL, WR, BL, BR, o[4], gflx, mflx);
    }
};

// ====================================================================
// Interface fluxes along j or k (D = 1, 2) with the stencil planes staged
// in LDS.  A workgroup owns 64 (i) x FH (along D) faces, FH = 4 or 8 (a
// launch-plan choice: the taller tile loads FH + 7 rows for FH face rows,
// fewer halo rows per face, with twice the waves per barrier); the rows
// of the 8-point stencil (faces' cells a-4 .. a+FH+2) and of the 4-point
// limiter stencil (a-2 .. a+FH) of one variable at a time go to LDS with
// coalesced loads, double-buffered so the next variable's rows are in
// flight while the current one is reconstructed.  Each row is loaded once
// per workgroup instead of once per stencil point.
// ====================================================================
struct FluxTile {
    int i0, i1, a0, a1, b0, b1;   // faces along i, along D, along the third axis
    int nti, nta, nb;
};

template <int D, int RM, bool BG, bool BORIS, int FH = BY, bool DMA = false>
struct k_flux_s {
    static constexpr int NT = BX * FH;                         // threads
    static constexpr int SR8 = FH + 7;                         // rows of an 8-point stage
    static constexpr int SR4 = FH + 3;                         // rows of a 4-point stage
    static constexpr int SRW = SR8 + SR4;                      // rows of one variable's stage
    static constexpr int SLD = (SRW * BX + NT - 1) / NT;       // loads per thread per stage
    __device__ __forceinline__ void operator()(const uint3 bi, Geo g, MParm m,
                                               FluxTile t, const double *__restrict__ H,
                                               const double *__restrict__ S,
                                               const double *__restrict__ met,
                                               const double *__restrict__ bg,
                                               double *__restrict__ gflx,
                                               double *__restrict__ mflx,
                                               double *__restrict__ Vf) const
    {
        __shared__ double vs[SR8][BX];
        __shared__ double buf[2][SRW][BX];
        const int w = (int)bi.x;
        const int per = (t.nb + 7) >> 3;
        const int L = (w & 7) * per + (w >> 3);
        if (L >= t.nb) return;                   // whole workgroup
        const int ti = L % t.nti;
        const int r = L / t.nti;
        const int ta = r % t.nta;
        const int bb = t.b0 + r / t.nta;
        const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + BX * ty;
        const int i0 = t.i0 + ti * BX, a0 = t.a0 + ta * FH;
        const int i = i0 + tx, a = a0 + ty;
        const bool act = i <= t.i1 && a <= t.a1;
        const long N3 = g.N3;
        const long s = STRIDE<D>(g);
        // point (i', a', bb) of the box's index space
        auto at = [&](int ii, int aa) -> long {
            return D == 1 ? IX(g, ii, aa, bb) : IX(g, ii, bb, aa);
        };
        // rows that exist: the stencils of the active faces reach a1 + 3
        const int aTop = t.a1 + 3;
        // stage loaders: element e of a stage -> (row, column)
        auto fetch = [&](const double *p8, const double *p4, double *reg) {
#pragma unroll
            for (int q = 0; q < SLD; q++) {
                const int e = tid + q * NT;
                const int row = e / BX, col = e % BX;
                const int ii = i0 + col;
                double v = 0.0;
                if (row < SRW && ii <= t.i1) {
                    const int aa = row < SR8 ? a0 - 4 + row : a0 - 2 + (row - SR8);
                    const double *p = row < SR8 ? p8 : p4;
                    if (aa <= aTop) v = p[at(ii, aa)];
                }
                reg[q] = v;
            }
        };
        auto store = [&](int bsel, const double *reg) {
#pragma unroll
            for (int q = 0; q < SLD; q++) {
                const int e = tid + q * NT;
                if (e < SRW * BX) buf[bsel][e / BX][e % BX] = reg[q];
            }
        };
        // DMA form: the stage's rows go straight from global memory to LDS
        // (global_load_lds, 16 B per lane): one instruction moves two rows,
        // lanes 0-31 the even row and 32-63 the odd one, into consecutive
        // LDS rows.  Columns past t.i1 read the tile's first pair and rows
        // past aTop the row aTop instead (valid cells; only the lanes of
        // faces outside the box read those LDS entries).
        auto dma = [&](const double *p8, const double *p4, int bsel) {
            const int wv = tid >> 6, ln = tid & 63;
            for (int q = wv; q < SRW / 2; q += NT / 64) {
                const int row = 2 * q + (ln >> 5);
                int ii = i0 + 2 * (ln & 31);
                if (ii > t.i1) ii = i0;
                int aa = row < SR8 ? a0 - 4 + row : a0 - 2 + (row - SR8);
                if (aa > aTop) aa = aTop;
                const double *src = (row < SR8 ? p8 : p4) + at(ii, aa);
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                                 (__attribute__((address_space(3))) void *)&buf[bsel][2 * q][0],
                                                 16, 0, 0);
            }
        };
        // volumes (8-point rows) and the first stage
        {
            const double *pv = met + M_VOL * N3;
            for (int e = tid; e < SR8 * BX; e += NT) {
                const int row = e / BX, col = e % BX;
                const int ii = i0 + col, aa = a0 - 4 + row;
                vs[row][col] = (ii <= t.i1 && aa <= aTop) ? pv[at(ii, aa)] : 0.0;
            }
        }
        const int nst = m.doMHD ? 8 : 5;
        const double *Wv = H + H_WH * N3, *Wl = H + H_WN * N3;
        const double *Bh = H + H_B * N3, *Bl = m.doLFM ? S + S_B * N3 : Bh;
        auto planes = [&](int st, const double *&p8, const double *&p4) {
            if (st < 5) { p8 = Wv + st * N3; p4 = Wl + st * N3; }
            else { p8 = Bh + (st - 5) * N3; p4 = Bl + (st - 5) * N3; }
        };
        double reg[DMA ? 1 : SLD];
        {
            const double *p8, *p4;
            planes(0, p8, p4);
            if (DMA) {
                dma(p8, p4, 0);
                __builtin_amdgcn_s_waitcnt(0);
            } else {
                fetch(p8, p4, reg);
                store(0, reg);
            }
        }
        __syncthreads();
        const double Vi = wsum8([&](int n) { return vs[ty + n][tx]; });
        const double rVi = 1.0 / Vi;
        double WL[5], WR[5], BL[3] = {0.0, 0.0, 0.0}, BR[3] = {0.0, 0.0, 0.0};
        // kept a loop: unrolled, the compiler hoists later stages' loads and
        // the register count doubles
#pragma unroll 1
        for (int st = 0; st < 8; st++) {
            if (st >= nst) break;
            const int cur = st & 1;
            if (!DMA && st + 1 < nst) {
                const double *p8, *p4;
                planes(st + 1, p8, p4);
                fetch(p8, p4, reg);
            }
            double q[8], nq[4], vq[8];
#pragma unroll
            for (int n = 0; n < 8; n++) q[n] = buf[cur][ty + n][tx];
#pragma unroll
            for (int n = 0; n < 4; n++) nq[n] = buf[cur][SR8 + ty + n][tx];
            if (DMA) {
                // this stage's LDS reads come first, then the next stage's
                // DMA, so that no LDS read waits behind the DMA in flight
                if (st >= 5) {
#pragma unroll
                    for (int n = 0; n < 8; n++) vq[n] = vs[ty + n][tx];
                }
                if (st + 1 < nst) {
                    const double *p8, *p4;
                    planes(st + 1, p8, p4);
                    dma(p8, p4, cur ^ 1);
                }
            }
            if (st < 5) {
                recon_lr_w<RM>(m.pdmb, rVi, nq, q, WL[st], WR[st]);
                // Mom2Face of this face from the same velocity stencil
                if (st >= 1 && st <= 3 && act) Vf[(st - 1) * N3 + at(i, a)] = mom2face(q, Vi, rVi);
            } else {
                // cell B is not volume-weighted in H: weight it here (LDS volumes)
#pragma unroll
                for (int n = 0; n < 8; n++) q[n] = (DMA ? vq[n] : vs[ty + n][tx]) * q[n];
                recon_lr_w<RM>(m.pdmb, rVi, nq, q, BL[st - 5], BR[st - 5]);
            }
            if (DMA) __builtin_amdgcn_s_waitcnt(0);
            else if (st + 1 < nst) store(cur ^ 1, reg);
            __syncthreads();
        }
        if (act) {
            const long c = at(i, a);
            const unsigned oc = (unsigned)(c * 8);
            face_flux_lr<D, BG, BORIS>(
                g, m, H, met, bg, [&](const double *p) { return ldp(p, oc); },
                [&](const double *p) { return p[c - s]; },
                [](double *, double *) {}, WL, WR, BL, BR, oc, gflx, mflx);
        }
    }
};

// ====================================================================
// CT update (E2Flux, fields.F90:18-51) + bFlux2Fld of the new state.
// Domain: all face points.  Faces outside is..ie+1 x js..je+1 x ks..ke+1
// keep their old flux; edges outside the active EMF boxes read as zero,
// as CalcElecField zeroes E before filling active edges (fields.F90:89).
// ====================================================================
__device__ __forceinline__ double e_at(const Geo &g, const EdgeBoxes &eb, const double *E, int d,
                                       int i, int j, int k, unsigned off)
{
    return in_box(eb.e[d], i, j, k) ? ldp(E + d * g.N3, off) : 0.0;
}

// New flux of face d at (i,j,k), whose byte offset is oc.
__device__ __forceinline__ double ct_face(const Geo &g, const EdgeBoxes &eb, const Box &ctb,
                                          const double *E, const double *mfOld, double dt,
                                          int d, int i, int j, int k, unsigned oc)
{
    const double old = ldp(mfOld + d * g.N3, oc);
    if (!in_box(ctb, i, j, k)) return old;
    const unsigned oI = oc + 8, oJ = oc + 8 * g.P, oK = oc + (unsigned)(8 * g.PP);
    double df;
    if (d == 0)
        df = e_at(g, eb, E, 2, i, j + 1, k, oJ) - e_at(g, eb, E, 2, i, j, k, oc)
           - e_at(g, eb, E, 1, i, j, k + 1, oK) + e_at(g, eb, E, 1, i, j, k, oc);
    else if (d == 1)
        df = e_at(g, eb, E, 0, i, j, k + 1, oK) - e_at(g, eb, E, 0, i, j, k, oc)
           - e_at(g, eb, E, 2, i + 1, j, k, oI) + e_at(g, eb, E, 2, i, j, k, oc);
    else
        df = e_at(g, eb, E, 1, i + 1, j, k, oI) - e_at(g, eb, E, 1, i, j, k, oc)
           - e_at(g, eb, E, 0, i, j + 1, k, oJ) + e_at(g, eb, E, 0, i, j, k, oc);
    return old - dt * df;
}

// ====================================================================
// Block min-reduction of positive doubles into clk->dtBits.
// ====================================================================
__device__ __forceinline__ void block_min_dt(double v, Clock *clk)
{
    __shared__ double red[512 / 64];   // up to 8 waves
    for (int off = 32; off > 0; off >>= 1) v = dmin(v, __shfl_xor(v, off, 64));
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double r = red[0];
        for (int w = 1; w < (int)(blockDim.x * blockDim.y) / 64; w++) r = dmin(r, red[w]);
        if (r > 0.0 && r < GAM_HUGE) atomicMin(&clk->dtBits, (unsigned long long)__double_as_longlong(r));
    }
}

// ====================================================================
// applyUpdate (mhdgroup.F90:79-258) in one pass over all face points:
//  * CT: E2Flux (fields.F90:18-51) on the point's three faces; faces
//    outside is..ie+1 x js..je+1 x ks..ke+1 keep the old flux, edges
//    outside the active EMF boxes read as zero (E is zeroed before the
//    active edges are filled, fields.F90:89);
//  * bFlux2Fld of the new state (the three upper faces are re-derived);
//  * Flux2Deltas (stress.F90:152-194), Reynolds with Model%dt (114-125)
//    and Maxwell/Boris (171-195, 299-441) on active cells; ghost cells copy
//    the old state (CopyState);
//  * CellDT of the new state on the CalcDT domain, min-reduced
//    (step.F90:91-101).
// Every thread reaches the block reduction.
// ====================================================================
template <bool BG, bool BORIS, int NTH = BX * BY>
struct k_update {
    static constexpr int NT = NTH;
    __device__ __forceinline__ void operator()(const uint3 bi, Geo g, MParm m,
                                               Tile ft, Box ctb, EdgeBoxes eb, Box ab, Box dtb,
                                               Clock *clk, const double *__restrict__ So,
                                               double *__restrict__ Sn,
                                               const double *__restrict__ gflx,
                                               const double *__restrict__ mflx,
                                               const double *__restrict__ met,
                                               const double *__restrict__ bg) const
    {
        int i, j, k;
        const bool live = tile_point(ft, (int)bi.x, i, j, k);
        double dtc = GAM_HUGE;
        const long N3 = g.N3;
        const long c = IX(g, i, j, k);
        const unsigned oc = (unsigned)(c * 8);    // byte offset of the point
        const bool cell = live && i <= g.ieg && j <= g.jeg && k <= g.keg;
        double Bn[3] = {0.0, 0.0, 0.0};
        if (live) {
            const double dt = clk->dtA;
            const double *E = Sn + S_E * N3;
            const double *mfo = So + S_MF * N3;
            double f[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                f[d] = m.doMHD ? ct_face(g, eb, ctb, E, mfo, dt, d, i, j, k, oc) : ldp(mfo + d * N3, oc);
                stp(Sn + (S_MF + d) * N3, oc, f[d]);
            }
            if (cell) {
                double fIp, fJp, fKp;
                if (m.doMHD) {
                    fIp = ct_face(g, eb, ctb, E, mfo, dt, 0, i + 1, j, k, oc + 8);
                    fJp = ct_face(g, eb, ctb, E, mfo, dt, 1, i, j + 1, k, oc + 8 * g.P);
                    fKp = ct_face(g, eb, ctb, E, mfo, dt, 2, i, j, k + 1, oc + (unsigned)(8 * g.PP));
                } else {
                    fIp = mfo[0 * N3 + c + 1]; fJp = mfo[1 * N3 + c + g.P]; fKp = mfo[2 * N3 + c + g.PP];
                }
                cell_b(g, met, c, f[0], f[1], f[2], fIp, fJp, fKp, Bn);
#pragma unroll
                for (int x = 0; x < 3; x++) stp(Sn + (S_B + x) * N3, oc, Bn[x]);
            }
        }
        if (cell) {
            double U[5];
#pragma unroll
            for (int v = 0; v < 5; v++) U[v] = ldp(So + (S_GAS + v) * N3, oc);
            if (in_box(ab, i, j, k)) {
                const double idV = 1.0 / ldp(met + M_VOL * N3, oc);
                const unsigned oI = oc + 8, oJ = oc + 8 * g.P, oK = oc + (unsigned)(8 * g.PP);
                const double dtM = clk->dt, dt = clk->dtA;
                double dH[5], oU[5];
#pragma unroll
                for (int v = 0; v < 5; v++) {
                    const double *gv = gflx + (long)v * N3;
                    dH[v] = (ldp(gv, oc) - ldp(gv, oI) + ldp(gv + 5 * N3, oc) - ldp(gv + 5 * N3, oJ)
                             + ldp(gv + 10 * N3, oc) - ldp(gv + 10 * N3, oK)) * idV;
                    oU[v] = U[v];
                    U[v] = U[v] + dtM * dH[v];
                }
                if (m.doMHD) {
                    double dM[3], B[3], oB[3];
#pragma unroll
                    for (int x = 0; x < 3; x++) {
                        const double *mv = mflx + (long)x * N3;
                        double r = (ldp(mv, oc) - ldp(mv, oI) + ldp(mv + 3 * N3, oc) - ldp(mv + 3 * N3, oJ)
                                    + ldp(mv + 6 * N3, oc) - ldp(mv + 6 * N3, oK)) * idV;
                        if (BG) r = r + ldp(bg + (B_DPB0 + x) * N3, oc);
                        dM[x] = r;
                        B[x] = Bn[x];
                        oB[x] = ldp(So + (S_B + x) * N3, oc);
                        if (BG) {
                            const double b0 = ldp(bg + (B_B0 + x) * N3, oc);
                            B[x] = B[x] + b0;
                            oB[x] = oB[x] + b0;
                        }
                    }
                    if (BORIS) {
                        // CellBoris (mhdgroup.F90:336-441), single fluid
                        double W2[5];
                        c2p_cell(m, U, W2);
                        const double D2 = dmax(W2[DEN], m.dFloor);
                        const double P2 = dmax(W2[PRS], m.pFloor);
                        const double D0 = dmax(oU[DEN], m.dFloor);
                        const double iD0 = 1.0 / D0, iD2 = 1.0 / D2;
                        double V0[3] = {oU[MOMX] * iD0, oU[MOMY] * iD0, oU[MOMZ] * iD0};
                        double B1[3] = {0.5 * (B[0] + oB[0]), 0.5 * (B[1] + oB[1]), 0.5 * (B[2] + oB[2])};
                        const double alpha = (B1[0] * B1[0] + B1[1] * B1[1] + B1[2] * B1[2]) * (iD2 * m.iCa2);
                        const double dRho = dt * dH[DEN];
                        double dPG[3] = {dt * dH[MOMX], dt * dH[MOMY], dt * dH[MOMZ]};
                        double dPM[3] = {dt * dM[0], dt * dM[1], dt * dM[2]};
                        const double pCon = 1.0 / (1.0 + alpha);
                        double dMom[3];
                        if (alpha > GAM_TINY) {
                            const double nb = sqrt(B1[0] * B1[0] + B1[1] * B1[1] + B1[2] * B1[2]);
                            double bh[3];
                            if (nb <= GAM_TINY) { bh[0] = bh[1] = bh[2] = 0.0; }
                            else { const double inb = 1.0 / nb; bh[0] = B1[0] * inb; bh[1] = B1[1] * inb; bh[2] = B1[2] * inb; }
                            double sGM[3] = {dPG[0] + dPM[0], dPG[1] + dPM[1], dPG[2] + dPM[2]};
                            const double pG = dPG[0] * bh[0] + dPG[1] * bh[1] + dPG[2] * bh[2];
                            const double pS = sGM[0] * bh[0] + sGM[1] * bh[1] + sGM[2] * bh[2];
                            const double pV = V0[0] * bh[0] + V0[1] * bh[1] + V0[2] * bh[2];
#pragma unroll
                            for (int x = 0; x < 3; x++)
                                dMom[x] = pG * bh[x] + pCon * (sGM[x] - pS * bh[x]) + alpha * pCon * dRho * (V0[x] - pV * bh[x]);
                        } else {
#pragma unroll
                            for (int x = 0; x < 3; x++) dMom[x] = dPG[x] + dPM[x];
                        }
                        double Wn[5];
                        Wn[DEN] = D2;
#pragma unroll
                        for (int x = 0; x < 3; x++) Wn[1 + x] = (D0 * V0[x] + dMom[x]) * iD2;
                        Wn[PRS] = P2;
                        p2c_cell(m, Wn, U);
                    } else {
                        // CellMaxwell (mhdgroup.F90:299-329)
                        const double d = dmax(U[DEN], m.dFloor);
                        double mx = U[MOMX], my = U[MOMY], mz = U[MOMZ];
                        const double id = 1.0 / d;
                        double ke = 0.5 * (mx * mx + my * my + mz * mz) * id;
                        const double p = dmax((m.gamma - 1) * (U[ENER] - ke), m.pFloor);
                        mx = mx + dt * dM[0];
                        my = my + dt * dM[1];
                        mz = mz + dt * dM[2];
                        ke = 0.5 * (mx * mx + my * my + mz * mz) * id;
                        U[DEN] = d; U[MOMX] = mx; U[MOMY] = my; U[MOMZ] = mz;
                        U[ENER] = ke + p * m.igm1;
                    }
                }
            }
#pragma unroll
            for (int v = 0; v < 5; v++) stp(Sn + (S_GAS + v) * N3, oc, U[v]);
            if (in_box(dtb, i, j, k)) {
                double bx = Bn[0], by = Bn[1], bz = Bn[2];
                if (BG) {
                    bx = bx + ldp(bg + (B_B0 + 0) * N3, oc);
                    by = by + ldp(bg + (B_B0 + 1) * N3, oc);
                    bz = bz + ldp(bg + (B_B0 + 2) * N3, oc);
                }
                const double dl = dmin(dmin(ldp(met + M_DI * N3, oc), ldp(met + M_DJ * N3, oc)), ldp(met + M_DK * N3, oc));
                dtc = cell_dt<BORIS>(m, U, bx, by, bz, dl);
                if (!(dtc == dtc)) dtc = GAM_HUGE;
            }
        }
        block_min_dt(dtc, clk);
    }
};

// CalcDT over the DT box on a state (standalone CalcDT, step.F90:58-101)
template <bool BG, bool BORIS>
struct k_calc_dt {
    __device__ __forceinline__ void operator()(const uint3 bi, Geo g, MParm m,
                                               Tile dtb, Clock *clk, const double *__restrict__ S,
                                               const double *__restrict__ met,
                                               const double *__restrict__ bg) const
    {
        int i, j, k;
        const bool live = tile_point(dtb, (int)bi.x, i, j, k);
        double dtc = GAM_HUGE;
        if (live) {
            const long N3 = g.N3, c = IX(g, i, j, k);
            double U[5];
            for (int v = 0; v < 5; v++) U[v] = S[(S_GAS + v) * N3 + c];
            double bx = S[(S_B + 0) * N3 + c], by = S[(S_B + 1) * N3 + c], bz = S[(S_B + 2) * N3 + c];
            if (BG) { bx = bx + bg[(B_B0 + 0) * N3 + c]; by = by + bg[(B_B0 + 1) * N3 + c]; bz = bz + bg[(B_B0 + 2) * N3 + c]; }
            const double dl = dmin(dmin(met[M_DI * N3 + c], met[M_DJ * N3 + c]), met[M_DK * N3 + c]);
            dtc = cell_dt<BORIS>(m, U, bx, by, bz, dl);
            if (!(dtc == dtc)) dtc = GAM_HUGE;
        }
        block_min_dt(dtc, clk);
    }
};

// ====================================================================
// Time bookkeeping and CalcDT checks (step.F90:103-166); one thread.
// ====================================================================
struct k_finalize {
    __device__ __forceinline__ void operator()(const uint3 bi, MParm m, Clock *clk, int mode) const
    {
        if (threadIdx.x != 0 || bi.x != 0) return;
        Clock c = *clk;
        if (mode & FIN_ADVANCE) {          // applyUpdate: oState%time=State%time; State%time+=dt
            c.ost = c.st;
            c.st = c.st + c.dtA;
        }
        if (mode & FIN_MODEL) {            // UpdateStateData (gamapp.F90:108-110)
            c.ts = c.ts + 1;
            c.t = c.t + c.dt;
        }
        if (mode & FIN_CALCDT) {
            const double dtMin = __longlong_as_double((long long)c.dtBits);
            int st = 0;
            if (dtMin == GAM_HUGE) st = 2;
            double dt = dtMin;
            if (c.ts > 0) {
                if (c.dt / dtMin >= 10) st = st > 1 ? st : 1;
                if ((m.dt0 > GAM_TINY) && (dtMin / m.dt0 <= m.limDT0)) st = st > 1 ? st : 1;
                if (dtMin <= GAM_TINY) st = 2;
            }
            if ((c.t + dt) > m.tFin) dt = dmax(m.tFin - c.t, GAM_TINY);
            if (m.dt0 > GAM_TINY && dt > (10.0 * m.dt0)) dt = 10.0 * m.dt0;
            c.dtRet = dt;
            c.lastStatus = st;
            c.status = c.status > st ? c.status : st;
            if (mode & FIN_SETDT) { c.dt = dt; c.dtA = dt; }
            if (mode & FIN_INIT) c.ost = c.st - c.dt;
        }
        c.dtBits = (unsigned long long)__double_as_longlong(GAM_HUGE);
        *clk = c;
    }
};

// ====================================================================
// bFlux2Fld (gridutils.F90:281-301) over all cells
// ====================================================================
struct k_bflux2fld {
    __device__ __forceinline__ void operator()(const uint3 bi, Geo g,
                                               Tile cb, Box skip, const double *__restrict__ mf,
                                               double *__restrict__ B,
                                               const double *__restrict__ met) const
    {
        TILE_IJK(cb);
        if (in_box(skip, i, j, k)) return;
        const long N3 = g.N3, c = IX(g, i, j, k);
        double b[3];
        cell_b(g, met, c, mf[c], mf[N3 + c], mf[2 * N3 + c], mf[c + 1], mf[N3 + c + g.P],
               mf[2 * N3 + c + g.PP], b);
        for (int x = 0; x < 3; x++) B[x * N3 + c] = b[x];
    }
};

// bFlux2Fld over the ghost shell only: up to six boxes (the i, j and k
// ghost slabs, disjoint), each cut into 64 x 4 (i, j) tiles; workgroup w
// finds its box from the running tile counts.
// A tile of BX*BY threads covers iw[q] points along i by BX*BY/iw[q] along
// j: 64 x 4 in the wide boxes, 4 x 64 in the 4-wide i-ghost slabs (every
// lane busy, 32 contiguous bytes per row).
struct GhostBoxes {
    Box b[6];
    int start[7];     // first tile of each box; start[n] = total
    int n0[6], n1[6], iw[6];
    int n;
};

struct k_bflux2fld_ghosts {
    __device__ __forceinline__ void operator()(const uint3 bi, Geo g,
                                               const GhostBoxes &gb, const double *__restrict__ mf,
                                               double *__restrict__ B,
                                               const double *__restrict__ met) const
    {
        const int w = (int)bi.x;
        int q = 0;
        while (q + 1 < gb.n && w >= gb.start[q + 1]) q++;
        const int L = w - gb.start[q];
        const Box &b = gb.b[q];
        const int t0 = L % gb.n0[q];
        const int r = L / gb.n0[q];
        const int tid = (int)(threadIdx.x + BX * threadIdx.y), iw = gb.iw[q];
        const int i = b.i0 + t0 * iw + tid % iw;
        const int j = b.j0 + (r % gb.n1[q]) * (BX * BY / iw) + tid / iw;
        const int k = b.k0 + r / gb.n1[q];
        if (i > b.i1 || j > b.j1 || k > b.k1) return;
        const long N3 = g.N3, c = IX(g, i, j, k);
        double bb[3];
        cell_b(g, met, c, mf[c], mf[N3 + c], mf[2 * N3 + c], mf[c + 1], mf[N3 + c + g.P],
               mf[2 * N3 + c + g.PP], bb);
        for (int x = 0; x < 3; x++) B[x * N3 + c] = bb[x];
    }
};

// ====================================================================
// Boundary conditions.  One thread per transverse ghost column with the n
// loop inside as in the reference (j and k sides), or per column and
// ghost layer (i sides, below).  Kernels of the six faces run in the
// reference order INI, OUTI, INJ, OUTJ, INK, OUTK.
// ====================================================================
__device__ __forceinline__ void copy_cell(double *S, long N3, long dst, long src)
{
    for (int v = 0; v < 5; v++) S[(S_GAS + v) * N3 + dst] = S[(S_GAS + v) * N3 + src];
}
__device__ __forceinline__ void copy_b(double *S, long N3, long dst, long src)
{
    for (int x = 0; x < 3; x++) S[(S_B + x) * N3 + dst] = S[(S_B + x) * N3 + src];
}
__device__ __forceinline__ bool is_cell(const Geo &g, int i, int j, int k)
{
    return i >= g.isg && i <= g.ieg && j >= g.jsg && j <= g.jeg && k >= g.ksg && k <= g.keg;
}

// Ghost layers n = 1..4: the i-side kernels (points (j, k) with 4-wide
// ghost slabs along i) run one layer per thread, the four layers of a point
// in adjacent lanes (a 64-lane block covers 16 points along j), so that one
// load reads 32 contiguous bytes along i rather than 8; the j- and k-side
// kernels (threads along i, coalesced) loop over n.  TA = the thread's
// transverse offset along the first transverse axis.
#define NL0(side) (((side) <= 1) ? 1 + (int)(threadIdx.x & 3) : 1)
#define NL1(side) (((side) <= 1) ? 1 + (int)(threadIdx.x & 3) : 4)
#define TA(side) (((side) <= 1) ? (int)(bi.x * 16 + (threadIdx.x >> 2)) : (int)(bi.x * blockDim.x + threadIdx.x))

// periodic (bcs.F90:279-450).  side: 0..5.  (a,b) = transverse indices
template <int SIDE>
struct k_bc_periodic {
    __device__ __forceinline__ void operator()(const uint3 bi, Geo g, Box t,
                                               double *__restrict__ S) const
    {
        const int a = t.i0 + TA(SIDE);
        const int b = t.j0 + bi.y;
        if (a > t.i1 || b > t.j1) return;
        const long N3 = g.N3;
        double *mf = S + S_MF * N3;
        for (int n = NL0(SIDE); n <= NL1(SIDE); n++) {
            if (SIDE == 0) {        // inner I: a=j, b=k
                const int ig = g.is - n, sr = g.ie - n + 1;
                const long d = IX(g, ig, a, b), s = IX(g, sr, a, b);
                if (is_cell(g, ig, a, b)) { copy_cell(S, N3, d, s); copy_b(S, N3, d, s); }
                for (int x = 0; x < 3; x++) mf[x * N3 + d] = mf[x * N3 + s];
            } else if (SIDE == 1) { // outer I
                const int ig = g.ie + n, sr = g.is + n - 1;
                const long d = IX(g, ig, a, b), s = IX(g, sr, a, b);
                if (is_cell(g, ig, a, b)) { copy_cell(S, N3, d, s); copy_b(S, N3, d, s); }
                mf[1 * N3 + d] = mf[1 * N3 + s];
                mf[2 * N3 + d] = mf[2 * N3 + s];
                mf[IX(g, ig + 1, a, b)] = mf[IX(g, g.is + n, a, b)];
            } else if (SIDE == 2) { // inner J: a=i, b=k
                const int jg = g.js - n, sr = g.je - n + 1;
                const long d = IX(g, a, jg, b), s = IX(g, a, sr, b);
                if (is_cell(g, a, jg, b)) { copy_cell(S, N3, d, s); copy_b(S, N3, d, s); }
                for (int x = 0; x < 3; x++) mf[x * N3 + d] = mf[x * N3 + s];
            } else if (SIDE == 3) { // outer J
                const int jg = g.je + n, sr = g.js + n - 1;
                const long d = IX(g, a, jg, b), s = IX(g, a, sr, b);
                if (is_cell(g, a, jg, b)) { copy_cell(S, N3, d, s); copy_b(S, N3, d, s); }
                mf[0 * N3 + d] = mf[0 * N3 + s];
                mf[2 * N3 + d] = mf[2 * N3 + s];
                mf[1 * N3 + IX(g, a, jg + 1, b)] = mf[1 * N3 + IX(g, a, g.js + n, b)];
            } else if (SIDE == 4) { // inner K: a=i, b=j
                const int kg = g.ks - n, sr = g.ke - n + 1;
                const long d = IX(g, a, b, kg), s = IX(g, a, b, sr);
                if (is_cell(g, a, b, kg)) { copy_cell(S, N3, d, s); copy_b(S, N3, d, s); }
                for (int x = 0; x < 3; x++) mf[x * N3 + d] = mf[x * N3 + s];
            } else {                // outer K
                const int kg = g.ke + n, sr = g.ks + n - 1;
                const long d = IX(g, a, b, kg), s = IX(g, a, b, sr);
                if (is_cell(g, a, b, kg)) { copy_cell(S, N3, d, s); copy_b(S, N3, d, s); }
                mf[0 * N3 + d] = mf[0 * N3 + s];
                mf[1 * N3 + d] = mf[1 * N3 + s];
                mf[2 * N3 + IX(g, a, b, kg + 1)] = mf[2 * N3 + IX(g, a, b, g.ks + n)];
            }
        }
    }
};

// zero-gradient (bcs.F90:799-930), sphere in/out (934-982)
template <int SIDE, int KIND>
struct k_bc_copy {
    __device__ __forceinline__ void operator()(const uint3 bi, Geo g, Box t,
                                               double *__restrict__ S,
                                               const double *__restrict__ met) const
    {
        const int a = t.i0 + TA(SIDE);
        const int b = t.j0 + bi.y;
        if (a > t.i1 || b > t.j1) return;
        const long N3 = g.N3;
        double *mf = S + S_MF * N3;
        if (SIDE == 0) {            // inner I: a=j, b=k
            const long s = IX(g, g.is, a, b);
            for (int n = NL0(SIDE); n <= NL1(SIDE); n++) {
                const long d = IX(g, g.is - n, a, b);
                copy_cell(S, N3, d, s);
                for (int x = 0; x < 3; x++) mf[x * N3 + d] = mf[x * N3 + s];
                if (KIND == 2) copy_b(S, N3, d, s);      // SphIn also copies Bxyz
            }
        } else if (SIDE == 1) {     // outer I
            const long s = IX(g, g.ie, a, b);
            if (KIND == 3) {        // SphOut
                for (int n = NL0(SIDE); n <= NL1(SIDE); n++) {
                    const int ig = g.ie + n;
                    const long d = IX(g, ig, a, b);
                    copy_cell(S, N3, d, s);
                    mf[IX(g, ig + 1, a, b)] = mf[s];
                    mf[N3 + d] = mf[N3 + s];
                    mf[2 * N3 + d] = mf[2 * N3 + s];
                    copy_b(S, N3, d, s);
                }
            } else {                // zeroGrad_obcI: fluxes of the last physical cell's B
                double B[3];
                cell_b(g, met, s, mf[s], mf[N3 + s], mf[2 * N3 + s], mf[s + 1], mf[N3 + s + g.P],
                       mf[2 * N3 + s + g.PP], B);
                for (int n = NL0(SIDE); n <= NL1(SIDE); n++) {
                    const int ig = g.ie + n;
                    const long d = IX(g, ig, a, b);
                    copy_cell(S, N3, d, s);
                    const long fi[3] = {IX(g, ig + 1, a, b), d, d};
                    for (int x = 0; x < 3; x++) {
                        const long f = fi[x];
                        const double *nrm = met + (long)(M_TF + 9 * x) * N3 + f;
                        mf[x * N3 + f] = met[(M_FACE + x) * N3 + f] * (B[0] * nrm[0] + B[1] * nrm[N3] + B[2] * nrm[2 * N3]);
                    }
                }
            }
        } else if (SIDE == 2) {     // inner J: a=i, b=k
            const long s = IX(g, a, g.js, b);
            for (int n = NL0(SIDE); n <= NL1(SIDE); n++) {
                const long d = IX(g, a, g.js - n, b);
                copy_cell(S, N3, d, s);
                for (int x = 0; x < 3; x++) mf[x * N3 + d] = mf[x * N3 + s];
            }
        } else if (SIDE == 3) {     // outer J
            const long s = IX(g, a, g.je, b);
            for (int n = NL0(SIDE); n <= NL1(SIDE); n++) {
                const long d = IX(g, a, g.je + n, b);
                copy_cell(S, N3, d, s);
                mf[d] = mf[s];
                mf[2 * N3 + d] = mf[2 * N3 + s];
                mf[N3 + IX(g, a, g.je + n + 1, b)] = mf[N3 + IX(g, a, g.je + 1, b)];
            }
        } else if (SIDE == 4) {     // inner K: a=i, b=j
            const long s = IX(g, a, b, g.ks);
            for (int n = NL0(SIDE); n <= NL1(SIDE); n++) {
                const long d = IX(g, a, b, g.ks - n);
                copy_cell(S, N3, d, s);
                for (int x = 0; x < 3; x++) mf[x * N3 + d] = mf[x * N3 + s];
            }
        } else {                    // outer K; the k+1 face takes the J flux (bcs.F90:925)
            const long s = IX(g, a, b, g.ke);
            for (int n = NL0(SIDE); n <= NL1(SIDE); n++) {
                const long d = IX(g, a, b, g.ke + n);
                copy_cell(S, N3, d, s);
                mf[d] = mf[s];
                mf[N3 + d] = mf[N3 + s];
                mf[2 * N3 + IX(g, a, b, g.ke + n + 1)] = mf[N3 + IX(g, a, b, g.ke + 1)];
            }
        }
    }
};

// ---------------------------------------------------------------- wind
#include "gam_wind.h"

// WindBC part 1 (wind.F90:226-335): ghost gas and blended face fluxes.
// Transverse: a = j in jsg..jeg+1, b = k in ksg..keg+1.
struct k_bc_wind {
    __device__ __forceinline__ void operator()(const uint3 bi, Geo g, MParm m, Box t,
                                               const Clock *clk, WindDev w, double *__restrict__ S,
                                               const double *__restrict__ met) const
    {
        const int j = t.i0 + TA(1);
        const int k = t.j0 + bi.y;
        if (j > t.i1 || k > t.j1) return;
        const long N3 = g.N3;
        double *mf = S + S_MF * N3;
        const double tm = clk->t;
        double vF[3];
        wind_front(w, tm, vF);
        const int ip = g.ie;
        for (int n = NL0(1); n <= NL1(1); n++) {
            const int ig = g.ie + n;
            if (k <= g.keg && j <= g.jeg) {
                const long cg = IX(g, ig, j, k), cp = IX(g, ip, j, k);
                const double xc[3] = {met[(M_XYZCC + 0) * N3 + cg], met[(M_XYZCC + 1) * N3 + cg], met[(M_XYZCC + 2) * N3 + cg]};
                double r, p, V[3], B[3];
                wind_at(w, vF, xc, tm, &r, &p, V, B);
                const double wSW = wind_wgt_v(xc, V), wMHD = 1.0 - wSW;
                double gs[5] = {0, 0, 0, 0, 0}, gi[5] = {0, 0, 0, 0, 0};
                if (wSW > GAM_TINY) {
                    gs[0] = r; gs[1] = V[0]; gs[2] = V[1]; gs[3] = V[2]; gs[4] = p;
                }
                if (wMHD > GAM_TINY) {
                    double U[5], Bi[3];
                    for (int v = 0; v < 5; v++) U[v] = S[(S_GAS + v) * N3 + cp];
                    for (int x = 0; x < 3; x++) Bi[x] = S[(S_B + x) * N3 + cp];
                    const long fp = IX(g, ip + 1, j, k);
                    const double nH[3] = {met[(M_TF + 0) * N3 + fp], met[(M_TF + 1) * N3 + fp], met[(M_TF + 2) * N3 + fp]};
                    wind_inner(m, U, Bi, nH, gi);
                }
                double gw[5], U[5];
                for (int v = 0; v < 5; v++) gw[v] = wSW * gs[v] + wMHD * gi[v];
                p2c_cell(m, gw, U);
                for (int v = 0; v < 5; v++) S[(S_GAS + v) * N3 + cg] = U[v];
            }
            const int fi[3] = {ig + 1, ig, ig};
            const int ii[3] = {ip + 1, ip, ip};
            for (int d = 0; d < 3; d++) {
                const long f = IX(g, fi[d], j, k);
                const double xf[3] = {met[(M_XFC + 0 + 3 * d) * N3 + f], met[(M_XFC + 1 + 3 * d) * N3 + f], met[(M_XFC + 2 + 3 * d) * N3 + f]};
                double r, p, V[3], B[3];
                wind_at(w, vF, xf, tm, &r, &p, V, B);
                const double wSW = wind_wgt_v(xf, V);
                const double *nrm = met + (long)(M_TF + 9 * d) * N3 + f;
                const double sw = met[(M_FACE + d) * N3 + f] * (nrm[0] * B[0] + nrm[N3] * B[1] + nrm[2 * N3] * B[2]);
                const double in = mf[d * N3 + IX(g, ii[d], j, k)];
                mf[d * N3 + f] = wSW * sw + (1 - wSW) * in;
            }
        }
    }
};

// WindBC part 2 (wind.F90:338-348): ghost Bxyz from the blended fluxes with
// the periodic k mapping of lfmIJKcc (ringutils.F90:482-518).
struct k_bc_wind_b {
    __device__ __forceinline__ void operator()(const uint3 bi, Geo g, Box t,
                                               double *__restrict__ S,
                                               const double *__restrict__ met) const
    {
        const int j = t.i0 + TA(1);
        const int k = t.j0 + bi.y;
        if (j > t.i1 || k > t.j1) return;
        const long N3 = g.N3;
        const double *mf = S + S_MF * N3;
        int kp = k;
        if (k < g.ks) kp = k + g.Nkp;
        else if (k > g.ke) kp = k - g.Nkp;
        for (int n = NL0(1); n <= NL1(1); n++) {
            const int ig = g.ie + n;
            const long s = IX(g, ig, j, kp);
            double B[3];
            cell_b(g, met, s, mf[s], mf[N3 + s], mf[2 * N3 + s], mf[s + 1], mf[N3 + s + g.P], mf[2 * N3 + s + g.PP], B);
            const long d = IX(g, ig, j, k);
            for (int x = 0; x < 3; x++) S[(S_B + x) * N3 + d] = B[x];
        }
    }
};

// DivB partial sums (gridutils.F90:717-770)
__global__ void k_divb(Geo g, Box ab, const double *__restrict__ mf, const double *__restrict__ bg,
                       int useBG, double *__restrict__ part)
{
    const int i = ab.i0 + blockIdx.x * BX + threadIdx.x;
    const int j = ab.j0 + blockIdx.y * BY + threadIdx.y;
    const int k = ab.k0 + blockIdx.z;
    double v = 0.0;
    if (i <= ab.i1 && j <= ab.j1) {
        const long N3 = g.N3, c = IX(g, i, j, k);
        double dFi = mf[c] - mf[c + 1], dFj = mf[N3 + c] - mf[N3 + c + g.P], dFk = mf[2 * N3 + c] - mf[2 * N3 + c + g.PP];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        if (useBG) {
            const double *b0 = bg + B_BFLUX0 * N3;
            a0 = b0[c] - b0[c + 1]; a1 = b0[N3 + c] - b0[N3 + c + g.P]; a2 = b0[2 * N3 + c] - b0[2 * N3 + c + g.PP];
        }
        v = fabs(dFi + dFj + dFk + a0 + a1 + a2);
    }
    __shared__ double red[BX * BY];
    const int tid = threadIdx.x + threadIdx.y * BX;
    red[tid] = v;
    __syncthreads();
    for (int s = BX * BY / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    if (tid == 0) part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = red[0];
}

// Layout conversion between Fortran host arrays and the padded device index
// space.  src/dst plane of an (ni,nj,nk) Fortran array starting at isg.
__global__ void k_repack_in(Geo g, int ni, int nj, int nk, const double *__restrict__ src,
                            double *__restrict__ dst)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tot = (long)ni * nj * nk;
    if (q >= tot) return;
    const int ii = (int)(q % ni);
    const long r = q / ni;
    const int jj = (int)(r % nj);
    const int kk = (int)(r / nj);
    dst[IX(g, g.isg + ii, g.jsg + jj, g.ksg + kk)] = src[q];
}

__global__ void k_repack_out(Geo g, int ni, int nj, int nk, const double *__restrict__ src,
                             double *__restrict__ dst)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tot = (long)ni * nj * nk;
    if (q >= tot) return;
    const int ii = (int)(q % ni);
    const long r = q / ni;
    const int jj = (int)(r % nj);
    const int kk = (int)(r / nj);
    dst[q] = src[IX(g, g.isg + ii, g.jsg + jj, g.ksg + kk)];
}

// Zero the E entries outside the active EMF boxes (CalcElecField zeroes E
// before filling the active edges, fields.F90:89).
__global__ void k_mask_e(Geo g, Box all, EdgeBoxes eb, double *__restrict__ E)
{
    THREAD_IJK(all);
    const long c = IX(g, i, j, k);
    for (int d = 0; d < 3; d++)
        if (!in_box(eb.e[d], i, j, k)) E[d * g.N3 + c] = 0.0;
}

// ====================================================================
// Host-side launchers
// ====================================================================
#define LAUNCH(kern, grid, block, stream, ...) hipLaunchKernelGGL(kern, grid, block, 0, stream, __VA_ARGS__)

Plan default_plan()
{
    Plan p;
    p.v[TS_PREDICT] = 3;
    p.v[TS_EMF_I] = 3;    // stencils along j and k
    p.v[TS_EMF_J] = 2;    // along i and k
    p.v[TS_EMF_K] = 1;    // along i and j
    p.v[TS_FLUX_I] = 1;
    p.v[TS_FLUX_J] = PLAN_LDSDMA;
    p.v[TS_FLUX_K] = PLAN_LDSDMA;
    p.v[TS_UPDATE] = 3;
    return p;
}

int plan_candidates(int slot, int *vals)
{
    int n = 0;
    if (slot < 0 || slot >= TS_COUNT) return 0;
    const int maxax = (slot == TS_FLUX_I) ? 2 : 3;
    // row order x chunk size (256, 512, 128 points per workgroup)
    for (int c = 0; c < 3; c++)
        for (int a = 1; a <= maxax; a++) vals[n++] = a + 8 * c;
    if (slot == TS_FLUX_J || slot == TS_FLUX_K) {
        vals[n++] = PLAN_LDS;
        vals[n++] = PLAN_LDS8;
        vals[n++] = PLAN_LDSDMA;
        vals[n++] = PLAN_LDS8DMA;
    }
    return n;
}

void launch_predict(hipStream_t s, int plan, const Geo &g, const MParm &m, const Box &fb, const Clock *clk,
                    const double *S, const double *O, const double *met, double *H)
{
    const int ch = plan_chunk(plan);
    const Tile t = tile_of(fb, plan_ax(plan), ch);
    if (ch == 512) launch_f<k_predict<512>>(tgrid(t), dim3(BX, 8), s, g, m, t, clk, S, O, met, H);
    else if (ch == 128) launch_f<k_predict<128>>(tgrid(t), dim3(BX, 2), s, g, m, t, clk, S, O, met, H);
    else launch_f<k_predict<>>(tgrid(t), dim3(BX, BY), s, g, m, t, clk, S, O, met, H);
}

void launch_vface_rim(hipStream_t s, const Geo &g, const double *H, const double *met, double *Vf)
{
    const Box bx[6] = {
        {g.is, g.ie + 1, g.js - 4, g.js - 1, g.ks, g.ke},      // I faces, j rim
        {g.is, g.ie + 1, g.je + 1, g.je + 4, g.ks, g.ke},
        {g.is, g.ie, g.js, g.je + 1, g.ks - 4, g.ks - 1},      // J faces, k rim
        {g.is, g.ie, g.js, g.je + 1, g.ke + 1, g.ke + 4},
        {g.is - 4, g.is - 1, g.js, g.je, g.ks, g.ke + 1},      // K faces, i rim
        {g.ie + 1, g.ie + 4, g.js, g.je, g.ks, g.ke + 1}};
    RimSet r;
    int tot = 0;
    for (int q = 0; q < 6; q++) {
        r.t[q] = tile_of(bx[q], q < 4 ? 2 : 1);
        r.start[q] = tot;
        tot += (int)tgrid(r.t[q]).x;
    }
    r.start[6] = tot;
    launch_f<k_vface_rim>(dim3((unsigned)tot), dim3(BX, BY), s, g, r, H, met, Vf);
}

template <int ED, int RM, bool BG, bool BO>
static void emf_t(hipStream_t s, int plan, const Geo &g, const MParm &m, const Box &b, const Clock *clk,
                  const double *H, const double *S, const double *Vf, const double *met,
                  const double *bg, double *E)
{
    const int ch = plan_chunk(plan);
    const Tile t = tile_of(b, plan_ax(plan), ch);
    if (ch == 512)
        launch_f<k_emf<ED, RM, BG, BO, 512>>(tgrid(t), dim3(BX, 8), s, g, m, t, clk, H, S, Vf, met, bg, E);
    else if (ch == 128)
        launch_f<k_emf<ED, RM, BG, BO, 128>>(tgrid(t), dim3(BX, 2), s, g, m, t, clk, H, S, Vf, met, bg, E);
    else
        launch_f<k_emf<ED, RM, BG, BO>>(tgrid(t), dim3(BX, BY), s, g, m, t, clk, H, S, Vf, met, bg, E);
}

template <int ED>
static void emf_d(hipStream_t s, int ax, const Geo &g, const MParm &m, const Box &b, const Clock *clk,
                  const double *H, const double *S, const double *Vf, const double *met,
                  const double *bg, double *E)
{
    const int sel = (m.rmeth ? 4 : 0) + (m.doBG ? 2 : 0) + (m.doBoris ? 1 : 0);
    switch (sel) {
    case 0: emf_t<ED, 0, false, false>(s, ax, g, m, b, clk, H, S, Vf, met, bg, E); break;
    case 1: emf_t<ED, 0, false, true>(s, ax, g, m, b, clk, H, S, Vf, met, bg, E); break;
    case 2: emf_t<ED, 0, true, false>(s, ax, g, m, b, clk, H, S, Vf, met, bg, E); break;
    case 3: emf_t<ED, 0, true, true>(s, ax, g, m, b, clk, H, S, Vf, met, bg, E); break;
    case 4: emf_t<ED, 1, false, false>(s, ax, g, m, b, clk, H, S, Vf, met, bg, E); break;
    case 5: emf_t<ED, 1, false, true>(s, ax, g, m, b, clk, H, S, Vf, met, bg, E); break;
    case 6: emf_t<ED, 1, true, false>(s, ax, g, m, b, clk, H, S, Vf, met, bg, E); break;
    default: emf_t<ED, 1, true, true>(s, ax, g, m, b, clk, H, S, Vf, met, bg, E); break;
    }
}

void launch_emf(hipStream_t s, int eD, int ax, const Geo &g, const MParm &m, const Box &b, const Clock *clk,
                const double *H, const double *S, const double *Vf, const double *met,
                const double *bg, double *E)
{
    if (eD == 0) emf_d<0>(s, ax, g, m, b, clk, H, S, Vf, met, bg, E);
    else if (eD == 1) emf_d<1>(s, ax, g, m, b, clk, H, S, Vf, met, bg, E);
    else emf_d<2>(s, ax, g, m, b, clk, H, S, Vf, met, bg, E);
}

template <int D, int RM, bool BG, bool BO>
static void flux_t(hipStream_t s, int form, const Geo &g, const MParm &m, const Box &b, const double *H,
                   const double *S, const double *met, const double *bg, double *gf, double *mf, double *vf)
{
    if constexpr (D != 0) {
        if (form >= PLAN_LDS && form <= PLAN_LDS8DMA) {
            const int fh = (form == PLAN_LDS8 || form == PLAN_LDS8DMA) ? 8 : BY;
            FluxTile t;
            t.i0 = b.i0;
            t.i1 = b.i1;
            if (D == 1) { t.a0 = b.j0; t.a1 = b.j1; t.b0 = b.k0; t.b1 = b.k1; }
            else { t.a0 = b.k0; t.a1 = b.k1; t.b0 = b.j0; t.b1 = b.j1; }
            t.nti = (t.i1 - t.i0 + BX) / BX;
            t.nta = (t.a1 - t.a0 + fh) / fh;
            t.nb = t.nti * t.nta * (t.b1 - t.b0 + 1);
            const dim3 grid((unsigned)(((t.nb + 7) / 8) * 8));
            if (form == PLAN_LDS8DMA)
                launch_f<k_flux_s<D, RM, BG, BO, 8, true>>(grid, dim3(BX, 8), s, g, m, t, H, S, met, bg, gf, mf, vf);
            else if (fh == 8)
                launch_f<k_flux_s<D, RM, BG, BO, 8>>(grid, dim3(BX, 8), s, g, m, t, H, S, met, bg, gf, mf, vf);
            else if (form == PLAN_LDSDMA)
                launch_f<k_flux_s<D, RM, BG, BO, BY, true>>(grid, dim3(BX, BY), s, g, m, t, H, S, met, bg, gf, mf,
                                                            vf);
            else
                launch_f<k_flux_s<D, RM, BG, BO>>(grid, dim3(BX, BY), s, g, m, t, H, S, met, bg, gf, mf, vf);
            return;
        }
    }
    const int ch = plan_chunk(form);
    const Tile t = tile_of(b, plan_ax(form), ch);
    if (ch == 512)
        launch_f<k_flux<D, RM, BG, BO, 512>>(tgrid(t), dim3(BX, 8), s, g, m, t, H, S, met, bg, gf, mf, vf);
    else if (ch == 128)
        launch_f<k_flux<D, RM, BG, BO, 128>>(tgrid(t), dim3(BX, 2), s, g, m, t, H, S, met, bg, gf, mf, vf);
    else
        launch_f<k_flux<D, RM, BG, BO>>(tgrid(t), dim3(BX, BY), s, g, m, t, H, S, met, bg, gf, mf, vf);
}

template <int D>
static void flux_d(hipStream_t s, int form, const Geo &g, const MParm &m, const Box &b, const double *H,
                   const double *S, const double *met, const double *bg, double *gf, double *mf, double *vf)
{
    const int sel = (m.rmeth ? 4 : 0) + (m.doBG ? 2 : 0) + (m.doBoris ? 1 : 0);
    switch (sel) {
    case 0: flux_t<D, 0, false, false>(s, form, g, m, b, H, S, met, bg, gf, mf, vf); break;
    case 1: flux_t<D, 0, false, true>(s, form, g, m, b, H, S, met, bg, gf, mf, vf); break;
    case 2: flux_t<D, 0, true, false>(s, form, g, m, b, H, S, met, bg, gf, mf, vf); break;
    case 3: flux_t<D, 0, true, true>(s, form, g, m, b, H, S, met, bg, gf, mf, vf); break;
    case 4: flux_t<D, 1, false, false>(s, form, g, m, b, H, S, met, bg, gf, mf, vf); break;
    case 5: flux_t<D, 1, false, true>(s, form, g, m, b, H, S, met, bg, gf, mf, vf); break;
    case 6: flux_t<D, 1, true, false>(s, form, g, m, b, H, S, met, bg, gf, mf, vf); break;
    default: flux_t<D, 1, true, true>(s, form, g, m, b, H, S, met, bg, gf, mf, vf); break;
    }
}

void launch_flux(hipStream_t s, int d, int form, const Geo &g, const MParm &m, const Box &b, const double *H,
                 const double *S, const double *met, const double *bg, double *gf, double *mf, double *vf)
{
    if (d == 0) flux_d<0>(s, form, g, m, b, H, S, met, bg, gf, mf, vf);
    else if (d == 1) flux_d<1>(s, form, g, m, b, H, S, met, bg, gf, mf, vf);
    else flux_d<2>(s, form, g, m, b, H, S, met, bg, gf, mf, vf);
}

template <bool BG, bool BO>
static void update_t(hipStream_t s, int plan, const Geo &g, const MParm &m, const Box &fb, const Box &ctb,
                     const EdgeBoxes &eb, const Box &ab, const Box &dtb, Clock *clk, const double *So,
                     double *Sn, const double *gf, const double *mf, const double *met, const double *bg)
{
    const int ch = plan_chunk(plan);
    const Tile t = tile_of(fb, plan_ax(plan), ch);
    if (ch == 512)
        launch_f<k_update<BG, BO, 512>>(tgrid(t), dim3(BX, 8), s, g, m, t, ctb, eb, ab, dtb, clk, So, Sn, gf, mf,
                                        met, bg);
    else if (ch == 128)
        launch_f<k_update<BG, BO, 128>>(tgrid(t), dim3(BX, 2), s, g, m, t, ctb, eb, ab, dtb, clk, So, Sn, gf, mf,
                                        met, bg);
    else
        launch_f<k_update<BG, BO>>(tgrid(t), dim3(BX, BY), s, g, m, t, ctb, eb, ab, dtb, clk, So, Sn, gf, mf, met,
                                   bg);
}

void launch_update(hipStream_t s, int plan, const Geo &g, const MParm &m, const Box &fb, const Box &ctb,
                   const Box *eboxes, const Box &ab, const Box &dtb, Clock *clk, const double *So,
                   double *Sn, const double *gf, const double *mf, const double *met, const double *bg)
{
    EdgeBoxes eb;
    for (int d = 0; d < 3; d++) eb.e[d] = eboxes[d];
    const int sel = (m.doBG ? 2 : 0) + (m.doBoris ? 1 : 0);
    switch (sel) {
    case 0: update_t<false, false>(s, plan, g, m, fb, ctb, eb, ab, dtb, clk, So, Sn, gf, mf, met, bg); break;
    case 1: update_t<false, true>(s, plan, g, m, fb, ctb, eb, ab, dtb, clk, So, Sn, gf, mf, met, bg); break;
    case 2: update_t<true, false>(s, plan, g, m, fb, ctb, eb, ab, dtb, clk, So, Sn, gf, mf, met, bg); break;
    default: update_t<true, true>(s, plan, g, m, fb, ctb, eb, ab, dtb, clk, So, Sn, gf, mf, met, bg); break;
    }
}

void launch_calc_dt(hipStream_t s, const Geo &g, const MParm &m, const Box &dtb, Clock *clk,
                    const double *S, const double *met, const double *bg)
{
    const int sel = (m.doBG ? 2 : 0) + (m.doBoris ? 1 : 0);
    const Tile t = tile_of(dtb, 1);
    switch (sel) {
    case 0: launch_f<k_calc_dt<false, false>>(tgrid(t), dim3(BX, BY), s, g, m, t, clk, S, met, bg); break;
    case 1: launch_f<k_calc_dt<false, true>>(tgrid(t), dim3(BX, BY), s, g, m, t, clk, S, met, bg); break;
    case 2: launch_f<k_calc_dt<true, false>>(tgrid(t), dim3(BX, BY), s, g, m, t, clk, S, met, bg); break;
    default: launch_f<k_calc_dt<true, true>>(tgrid(t), dim3(BX, BY), s, g, m, t, clk, S, met, bg); break;
    }
}

void launch_finalize(hipStream_t s, const MParm &m, Clock *clk, int mode)
{
    launch_f<k_finalize>(dim3(1), dim3(64), s, m, clk, mode);
}

void launch_bflux2fld(hipStream_t s, const Geo &g, const Box &cb, const Box &skip, const double *mf,
                      double *B, const double *met)
{
    const Tile t = tile_of(cb, 1);
    launch_f<k_bflux2fld>(tgrid(t), dim3(BX, BY), s, g, t, skip, mf, B, met);
}

void launch_bflux2fld_ghosts(hipStream_t s, const Geo &g, const double *mf, double *B, const double *met)
{
    GhostBoxes gb;
    const Box bx[6] = {
        {g.isg, g.ieg, g.jsg, g.jeg, g.ksg, g.ks - 1},      // k ghosts, all i, j
        {g.isg, g.ieg, g.jsg, g.jeg, g.ke + 1, g.keg},
        {g.isg, g.ieg, g.jsg, g.js - 1, g.ks, g.ke},        // j ghosts, active k
        {g.isg, g.ieg, g.je + 1, g.jeg, g.ks, g.ke},
        {g.isg, g.is - 1, g.js, g.je, g.ks, g.ke},          // i ghosts, active j, k
        {g.ie + 1, g.ieg, g.js, g.je, g.ks, g.ke}};
    int tot = 0;
    for (int q = 0; q < 6; q++) {
        gb.b[q] = bx[q];
        gb.start[q] = tot;
        gb.iw[q] = (q >= 4) ? 4 : BX;
        const int jw = BX * BY / gb.iw[q];
        gb.n0[q] = (bx[q].i1 - bx[q].i0 + gb.iw[q]) / gb.iw[q];
        gb.n1[q] = (bx[q].j1 - bx[q].j0 + jw) / jw;
        tot += gb.n0[q] * gb.n1[q] * (bx[q].k1 - bx[q].k0 + 1);
    }
    gb.start[6] = tot;
    gb.n = 6;
    launch_f<k_bflux2fld_ghosts>(dim3((unsigned)tot), dim3(BX, BY), s, g, gb, mf, B, met);
}

static inline dim3 grid2(const Box &t) { return dim3((unsigned)((t.i1 - t.i0 + 64) / 64), (unsigned)(t.j1 - t.j0 + 1)); }

// Both periodic sides of one axis in one launch: each reads only interior
// cells and faces and writes its own ghosts, so the two commute.  Blocks
// with z = 0 run the low side, z = 1 the high side.
template <int AX>
struct k_bc_periodic_pair {
    __device__ __forceinline__ void operator()(const uint3 bi, Geo g, Box t, double *__restrict__ S) const
    {
        if (bi.z == 0) k_bc_periodic<2 * AX>{}(bi, g, t, S);
        else k_bc_periodic<2 * AX + 1>{}(make_uint3(bi.x, bi.y, 0), g, t, S);
    }
};

void launch_bc_periodic_pair(hipStream_t s, int ax, const Geo &g, double *S)
{
    Box t;   // the transverse boxes of launch_bc's periodic sides
    if (ax == 0) { t.i0 = g.jsg; t.i1 = g.jeg + 1; t.j0 = g.ksg; t.j1 = g.keg + 1; }
    else if (ax == 1) { t.i0 = g.isg; t.i1 = g.ieg + 1; t.j0 = g.ksg; t.j1 = g.keg + 1; }
    else { t.i0 = g.isg; t.i1 = g.ieg + 1; t.j0 = g.jsg; t.j1 = g.jeg + 1; }
    t.k0 = t.k1 = 0;
    dim3 gr = grid2(t), bl(64);
    if (ax == 0) gr.x = (unsigned)((t.i1 - t.i0 + 16) / 16);   // 16 points x 4 layers per block
    gr.z = 2;
    if (ax == 0) launch_f<k_bc_periodic_pair<0>>(gr, bl, s, g, t, S);
    else if (ax == 1) launch_f<k_bc_periodic_pair<1>>(gr, bl, s, g, t, S);
    else launch_f<k_bc_periodic_pair<2>>(gr, bl, s, g, t, S);
}

void launch_bc(hipStream_t s, int side, int kind, const Geo &g, const MParm &m, const Clock *clk,
               const WindDev *w, double *S, const double *met, bool ghostB)
{
    // Transverse boxes (a,b) per side as in the reference loops
    Box t;
    const bool pi = (kind == 0);
    if (side <= 1) { t.i0 = g.jsg; t.i1 = g.jeg + (pi ? 1 : 0); t.j0 = g.ksg; t.j1 = g.keg + (pi ? 1 : 0); }
    else if (side <= 3) { t.i0 = g.isg; t.i1 = g.ieg + (pi ? 1 : 0); t.j0 = g.ksg; t.j1 = g.keg + (pi ? 1 : 0); }
    else { t.i0 = g.isg; t.i1 = g.ieg + (pi ? 1 : 0); t.j0 = g.jsg; t.j1 = g.jeg + (pi ? 1 : 0); }
    t.k0 = t.k1 = 0;
    dim3 gr = grid2(t), bl(64);
    if (side <= 1) gr.x = (unsigned)((t.i1 - t.i0 + 16) / 16);   // 16 points x 4 layers per block
    if (kind == 0) {
        switch (side) {
        case 0: launch_f<k_bc_periodic<0>>(gr, bl, s, g, t, S); break;
        case 1: launch_f<k_bc_periodic<1>>(gr, bl, s, g, t, S); break;
        case 2: launch_f<k_bc_periodic<2>>(gr, bl, s, g, t, S); break;
        case 3: launch_f<k_bc_periodic<3>>(gr, bl, s, g, t, S); break;
        case 4: launch_f<k_bc_periodic<4>>(gr, bl, s, g, t, S); break;
        default: launch_f<k_bc_periodic<5>>(gr, bl, s, g, t, S); break;
        }
    } else if (kind == 1) {
        switch (side) {
        case 0: launch_f<k_bc_copy<0, 1>>(gr, bl, s, g, t, S, met); break;
        case 1: launch_f<k_bc_copy<1, 1>>(gr, bl, s, g, t, S, met); break;
        case 2: launch_f<k_bc_copy<2, 1>>(gr, bl, s, g, t, S, met); break;
        case 3: launch_f<k_bc_copy<3, 1>>(gr, bl, s, g, t, S, met); break;
        case 4: launch_f<k_bc_copy<4, 1>>(gr, bl, s, g, t, S, met); break;
        default: launch_f<k_bc_copy<5, 1>>(gr, bl, s, g, t, S, met); break;
        }
    } else if (kind == 2) {
        launch_f<k_bc_copy<0, 2>>(gr, bl, s, g, t, S, met);
    } else if (kind == 3) {
        launch_f<k_bc_copy<1, 3>>(gr, bl, s, g, t, S, met);
    } else if (kind == 4) {
        Box tw = t;
        tw.i1 = g.jeg + 1; tw.j1 = g.keg + 1;
        dim3 gw = grid2(tw);
        gw.x = (unsigned)((tw.i1 - tw.i0 + 16) / 16);
        launch_f<k_bc_wind>(gw, bl, s, g, m, tw, clk, *w, S, met);
        // ghost-cell B of the wind BC (wind.F90:338-348); skipped when a
        // bFlux2Fld over every ghost cell follows, which overwrites it
        if (ghostB) launch_f<k_bc_wind_b>(gr, bl, s, g, t, S, met);
    }
}

void launch_divb(hipStream_t s, const Geo &g, const Box &ab, const double *mf, const double *bg,
                 int useBG, double *part, int *nparts)
{
    dim3 gr = grid_of(ab);
    *nparts = (int)(gr.x * gr.y * gr.z);
    LAUNCH(k_divb, gr, dim3(BX, BY), s, g, ab, mf, bg, useBG, part);
}

void launch_repack(hipStream_t s, int dir_in, const Geo &g, int ni, int nj, int nk,
                   const double *src, double *dst)
{
    const long tot = (long)ni * nj * nk;
    dim3 gr((unsigned)((tot + 255) / 256));
    if (dir_in) LAUNCH(k_repack_in, gr, dim3(256), s, g, ni, nj, nk, src, dst);
    else LAUNCH(k_repack_out, gr, dim3(256), s, g, ni, nj, nk, src, dst);
}

void launch_mask_e(hipStream_t s, const Geo &g, const Box &all, const Box *eboxes, double *E)
{
    EdgeBoxes eb;
    for (int d = 0; d < 3; d++) eb.e[d] = eboxes[d];
    LAUNCH(k_mask_e, grid_of(all), dim3(BX, BY), s, g, all, eb, E);
}
