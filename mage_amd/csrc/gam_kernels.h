// gam_kernels.h -- host-side launchers of the gfx950 kernels.
#pragma once
#include "gam_common.h"
#include <string.h>
#include <vector>

// Argument record of one kernel launch: the arguments in order, as a
// nested aggregate (a0, then the record of the rest).  The host writes it
// to device memory for batched launches; layout is the same on both sides.
template <class... A>
struct Pack {};
template <class A0, class... A>
struct Pack<A0, A...> {
    A0 a0;
    Pack<A...> r;
};
__host__ __device__ inline Pack<> make_pack() { return Pack<>{}; }
template <class A0, class... A>
__host__ __device__ inline Pack<A0, A...> make_pack(A0 a0, A... a)
{
    return Pack<A0, A...>{a0, make_pack(a...)};
}

// Batched launch over the members of an ensemble: member m owns workgroups
// wg0[m] .. wg0[m+1]-1 (wg0[m] a multiple of 8) of which the first
// gx*gy*gz run its own grid.  In device memory the header is followed by
// an unsigned short per run of 8 workgroups: the member that owns it.
#define ENS_MAX 256
struct EnsHdr {
    unsigned n;
    unsigned wg0[ENS_MAX + 1];
    unsigned gx[ENS_MAX], gy[ENS_MAX], gz[ENS_MAX];
};

// One recorded launch of a step kernel: the function that issues the
// batched form for this kernel and argument types, the member's own grid
// and block, and the argument record bytes.
struct EnsLaunch {
    void (*launch)(hipStream_t s, const void *hdr, const void *rec, unsigned nwg, dim3 block);
    dim3 grid, block;
    std::vector<unsigned char> rec;
    hipStream_t stream;   // the stream the launch or marker was issued to
    int mark;             // 0 launch; 1 event ev recorded on stream; 2 stream waits for event ev
    int ev;
};
// While r is set, the step kernels that this thread would launch are
// appended to *r instead (nothing is issued); nullptr restores launching.
void set_launch_recorder(std::vector<EnsLaunch> *r);
// Stream order of a branching step: hipEventRecord(e, s) and
// hipStreamWaitEvent(s, e) at once, or, while a recorder is set, appended
// to it as markers naming the event by its index ev.
void order_record(hipStream_t s, hipEvent_t e, int ev);
void order_wait(hipStream_t s, hipEvent_t e, int ev);

// Launch plan: the shape of every tunable launch of the step.  Values
// 1, 2, 3 are the Tile row orders (ax1: j fastest, k fastest, j-slabs);
// PLAN_LDS and PLAN_LDS8 select the LDS-staged form of the j/k interface
// fluxes (a 64 x 4 or 64 x 8 face tile with its stencil halo rows in LDS);
// PLAN_LDSDMA and PLAN_LDS8DMA the 64 x 4 and 64 x 8 tiles whose rows reach
// LDS by global_load_lds.
enum { TS_PREDICT = 0, TS_EMF_I, TS_EMF_J, TS_EMF_K, TS_FLUX_I, TS_FLUX_J, TS_FLUX_K, TS_UPDATE,
       TS_COUNT };
#define PLAN_LDS 4
#define PLAN_LDS8 5
#define PLAN_LDSDMA 6
#define PLAN_LDS8DMA 7
struct Plan {
    int v[TS_COUNT];
};
Plan default_plan();
// candidate values of a slot; returns their number
int plan_candidates(int slot, int *vals);

// k_finalize modes
enum {
    FIN_ADVANCE = 1,   // State/oState times after applyUpdate
    FIN_MODEL = 2,     // Model%ts += 1, Model%t += Model%dt
    FIN_CALCDT = 4,    // evaluate CalcDT from the reduction
    FIN_SETDT = 8,     // Model%dt = CalcDT result
    FIN_INIT = 16      // oState%time = State%time - dt (init.F90:238-244)
};

// Sets the message gam_last_error() returns ("msg: detail") and returns code.
int gam_fail_(int code, const char *msg, const char *detail);

void launch_predict(hipStream_t s, int ax, const Geo &g, const MParm &m, const Box &fb, const Clock *clk,
                    const double *S, const double *O, const double *met, double *H);
// Mom2Face on the faces the EMF stencils read past the interface-flux faces
// (the flux kernels write the rest): Vf planes [3*dT1 + comp]
void launch_vface_rim(hipStream_t s, const Geo &g, const double *H, const double *met, double *Vf);
void launch_emf(hipStream_t s, int eD, int ax, const Geo &g, const MParm &m, const Box &b, const Clock *clk,
                const double *H, const double *S, const double *Vf, const double *met,
                const double *bg, double *E);
// Interface fluxes of direction d, and Mom2Face of the same faces into
// vf's three planes of direction d
void launch_flux(hipStream_t s, int d, int form, const Geo &g, const MParm &m, const Box &b, const double *H,
                 const double *S, const double *met, const double *bg, double *gf, double *mf, double *vf);
void launch_update(hipStream_t s, int ax, const Geo &g, const MParm &m, const Box &fb, const Box &ctb,
                   const Box *eboxes, const Box &ab, const Box &dtb, Clock *clk, const double *So,
                   double *Sn, const double *gf, const double *mf, const double *met, const double *bg);
void launch_calc_dt(hipStream_t s, const Geo &g, const MParm &m, const Box &dtb, Clock *clk,
                    const double *S, const double *met, const double *bg);
void launch_finalize(hipStream_t s, const MParm &m, Clock *clk, int mode);
void launch_bflux2fld(hipStream_t s, const Geo &g, const Box &cb, const Box &skip, const double *mf,
                      double *B, const double *met);
void launch_bflux2fld_ghosts(hipStream_t s, const Geo &g, const double *mf, double *B, const double *met);
// ghostB = false skips ghost-cell B that a following ghost bFlux2Fld rewrites
// both sides of axis ax (0 = I, 1 = J, 2 = K) when both are periodic
void launch_bc_periodic_pair(hipStream_t s, int ax, const Geo &g, double *S);
void launch_bc(hipStream_t s, int side, int kind, const Geo &g, const MParm &m, const Clock *clk,
               const WindDev *w, double *S, const double *met, bool ghostB);
void launch_divb(hipStream_t s, const Geo &g, const Box &ab, const double *mf, const double *bg,
                 int useBG, double *part, int *nparts);
void launch_repack(hipStream_t s, int dir_in, const Geo &g, int ni, int nj, int nk,
                   const double *src, double *dst);
void launch_mask_e(hipStream_t s, const Geo &g, const Box &all, const Box *eboxes, double *E);

// Grid pre-compute (gam_grid.hip)
void launch_grid_from_corners(hipStream_t s, const Geo &g, double *met);
void launch_dipole_b0(hipStream_t s, const Geo &g, double M0, const double *met, double *bg,
                      double *scratch9);

// Adjoint of the hydrodynamic step (gam_adj.hip)
void launch_adj_objective(hipStream_t s, const Geo &g, const MParm &m, const double *S,
                          const double *wgt, double *aS, double *part, int *nparts);
void launch_bc_adj(hipStream_t s, int side, int kind, const Geo &g, double *aS);
// pdt (may be null): the dt-sensitivity plane (d/d(dt) shares, see
// launch_dt_fold); pw: the AB2-weight plane
void launch_update_adj(hipStream_t s, const Geo &g, const Box &ab, const Clock *clk, const double *aSn,
                       double *aSo, double *ag, const double *met, const double *gflx, double *pdt);
void launch_flux_adj(hipStream_t s, int d, const Geo &g, const MParm &m, const Box &fb, const double *H,
                     const double *met, const double *ag, double *aWH, double *aWN);
void launch_predict_adj(hipStream_t s, const Geo &g, const MParm &m, const Clock *clk, const double *S,
                        const double *O, const double *met, const double *aWH, const double *aWN,
                        double *aS, double *aO, double *pw);
void launch_set_dt(hipStream_t s, Clock *clk, const Clock *rec);
// dt = dtA = dts[n] on the device clock (replays with a given dt sequence)
void launch_set_dt_val(hipStream_t s, Clock *clk, const double *dts, int n);
// Reduce one adjoint step's dt planes (pdt, pw over the face box) into
// mu[n] (and mu[n-1] through the AB2 weight's odt); part holds 2 doubles
// per 64 x 4 block of the face box (dt_part_blocks).
long dt_part_blocks(const Geo &g);
void launch_dt_fold(hipStream_t s, const Geo &g, const double *pdt, const double *pw, double *part,
                    const Clock *clk, int n, double *mu);
// CalcDT's argmin over the DT box of S (the state the update has just
// written, before the boundary pass), recorded with its values as
// DT_REC doubles {cell, gas 5, total cell B 3, dl, valid} when its CellDT
// is the dt clk holds (valid = 0 where tFin or 10 dt0 capped it); part: 2
// doubles per 64 x 4 block of dtb (dt_record_blocks)
#define DT_REC 12
long dt_record_blocks(const Geo &g, const Box &dtb);
void launch_dt_record(hipStream_t s, const Geo &g, const MParm &m, const Box &dtb, const double *S,
                      const double *met, const double *bg, const Clock *clk, double *part, double *rec);
// CalcDT transpose: lam += mu * d(CellDT)/d(gas, cell B) at rec's cell
void launch_dt_inject(hipStream_t s, const Geo &g, const MParm &m, const double *rec, const double *mu,
                      double *lam);
void launch_cellb_adj(hipStream_t s, const Geo &g, int all, double *aS, const double *met);
void launch_bc_periodic_adj(hipStream_t s, int side, const Geo &g, double *aS);
void launch_bc_mhd_adj(hipStream_t s, int side, int kind, const Geo &g, double *aS, const double *met);
// transpose of the inflow BC; adds the series adjoint into gQ (n,5), gB (n,3),
// each WIND_COPIES consecutive copies of those arrays (the caller sums them)
#define WIND_COPIES 32
void launch_bc_wind_adj(hipStream_t s, const Geo &g, const MParm &m, const Clock *clk, const WindDev &w,
                        const double *S, double *aS, const double *met, double *gQ, double *gB);
void launch_update_adj_mhd(hipStream_t s, const Geo &g, const MParm &m, const Box &ab, const Box &ctb,
                           const Box *eboxes, const Clock *clk, const double *So, const double *Sn1,
                           const double *gflx, const double *mflx, const double *met, const double *bg,
                           double *aSn, double *aSo, double *aUr, double *adM, double *ag, double *amf,
                           double *aE, double *pdt);
// pass 1 of the flux transpose: per face, the stencil scalars and limiter
// adjoints of its 8 reconstructions into sc (48 planes: per variable the
// pair (sl, sr) interleaved over two planes, then the four limiter
// adjoints), aBn into aHMF, the Boris nuke term into aHG
void launch_flux_adj_mhd(hipStream_t s, int d, const Geo &g, const MParm &m, const Box &fb, const double *H,
                         const double *S, const double *met, const double *bg, const double *ag,
                         const double *amf, double *sc, double *aHMF, double *aHG);
// pass 2 (all three directions' sc, 48 planes each) fused with the Mom2Face
// transpose: writes aWH, aWN, aHB of every cell
void launch_recon_adj_gather(hipStream_t s, const Geo &g, const MParm &m, const double *sc, const double *rv,
                             const double *aVf, const double *met, double *aWH, double *aWN, double *aHB,
                             double *aSSB);
// pass 1 of the edge EMF transpose: 16 planes per edge direction into sc
void launch_emf_adj(hipStream_t s, int ed, const Geo &g, const MParm &m, const Box &eb, const Clock *clk,
                    const double *H, const double *S, const double *Vf, const double *met, const double *bg,
                    const double *aE, double *sc, double *pdt);
// pass 2: writes aVf, aHMF and aHG at every point (adds the limiter parts
// to aSSMF with the LFM limiter)
void launch_emf_adj_gather(hipStream_t s, const Geo &g, const MParm &m, const Box *eboxes, const double *sc,
                           const double *met, double *aVf, double *aHMF, double *aHG, double *aSSMF);
void launch_vface_rv(hipStream_t s, const Geo &g, const double *met, double *rv);
void launch_predict_adj_mhd(hipStream_t s, const Geo &g, const MParm &m, const Clock *clk, const double *S,
                            const double *O, const double *met, const double *aHG, const double *aWH,
                            const double *aWN, double *aHB, double *aHMF, double *aS, double *aO, double *pw);
