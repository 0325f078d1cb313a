// gam_noalloc.hip -- total |div B| of face fluxes held in device memory,
// with no device allocation.
//
// gam_divb (gam_api.hip) reduces the State of a context through per-block
// partial sums in a buffer the context allocated.  gam_divb_noalloc takes
// the face fluxes (and optionally the background's face fluxes) as device
// arrays in the reference's Fortran layout, needs no context, and keeps its
// partial sums in `scratch`, a file-scope device array of fixed size that
// every GPU holds for the life of the process.  It is for callers that
// cannot allocate: code under HIP graph capture, tight loops, harnesses
// that forbid hipMalloc.
#include "gamhip.h"
#include "gam_kernels.h"
#include <hip/hip_runtime.h>
#include <mutex>

#define NA_BLOCKS 2048
#define NA_THREADS 256
#define NA_MAX_DEVICES 64

// per-block partial sums of one call: NA_BLOCKS doubles (16 KiB) per GPU
static __device__ __align__(256) double scratch[NA_BLOCKS];

// `scratch` of each device, as a device address (the array's name is a
// host-side stand-in in host code)
static double *g_scratch[NA_MAX_DEVICES];
static std::mutex g_scratch_mx;

static int scratch_of_current_device(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return gam_fail_(GAM_E_HIP, "gam_divb_noalloc: hipGetDevice failed", hipGetErrorString(e));
    if (dev < 0 || dev >= NA_MAX_DEVICES) return gam_fail_(GAM_E_ARG, "gam_divb_noalloc: device index out of range", "");
    std::lock_guard<std::mutex> lk(g_scratch_mx);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return gam_fail_(GAM_E_HIP, "gam_divb_noalloc: hipGetSymbolAddress(scratch) failed",
                             hipGetErrorString(e));
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return GAM_OK;
}

// Block sum of v over 256 threads in a fixed order (tree in LDS).
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double red[NA_THREADS];
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = NA_THREADS / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    return red[0];
}

// |div B| of every active cell (DivB, gridutils.F90:717-770), grid-stride
// over the cells; block b writes part[b] (every block writes its entry, so
// nothing needs zeroing between calls).  mf: (Ni+1,Nj+1,Nk+1,3) face
// fluxes; b0: (Ni,Nj,Nk,3) background face fluxes or null.
__global__ __launch_bounds__(NA_THREADS) void k_divb_noalloc(int Nip, int Njp, int Nkp, int nG,
                                                             const double *__restrict__ mf,
                                                             const double *__restrict__ b0,
                                                             double *__restrict__ part)
{
    const long Ni = Nip + 2 * nG, Nj = Njp + 2 * nG, Nk = Nkp + 2 * nG;
    const long fP = Ni + 1, fPP = (Ni + 1) * (Nj + 1), fN = fPP * (Nk + 1);
    const long cP = Ni, cPP = Ni * Nj, cN = cPP * Nk;
    const long ncell = (long)Nip * Njp * Nkp;
    double acc = 0.0;
    for (long n = (long)blockIdx.x * NA_THREADS + threadIdx.x; n < ncell; n += (long)NA_BLOCKS * NA_THREADS) {
        const long i = n % Nip + nG, r = n / Nip;
        const long j = r % Njp + nG, k = r / Njp + nG;
        const long f = i + fP * j + fPP * k;
        const double dFi = mf[f] - mf[f + 1];
        const double dFj = mf[fN + f] - mf[fN + f + fP];
        const double dFk = mf[2 * fN + f] - mf[2 * fN + f + fPP];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        if (b0) {
            const long c = i + cP * j + cPP * k;
            a0 = b0[c] - b0[c + 1];
            a1 = b0[cN + c] - b0[cN + c + cP];
            a2 = b0[2 * cN + c] - b0[2 * cN + c + cPP];
        }
        acc += fabs(dFi + dFj + dFk + a0 + a1 + a2);
    }
    const double s = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// One block: the NA_BLOCKS partial sums into *out (fixed order).
__global__ __launch_bounds__(NA_THREADS) void k_sum_noalloc(const double *__restrict__ part,
                                                            double *__restrict__ out)
{
    double v = 0.0;
    for (int q = threadIdx.x; q < NA_BLOCKS; q += NA_THREADS) v += part[q];
    const double s = block_sum(v);
    if (threadIdx.x == 0) *out = s;
}

extern "C" int gam_divb_noalloc(const gam_dims *dims, const double *magFlux, const double *bFlux0,
                                void *stream, double *totDivB)
{
    if (!dims || !magFlux || !totDivB) return gam_fail_(GAM_E_ARG, "gam_divb_noalloc: NULL argument", "");
    const long Nip = dims->Nip, Njp = dims->Njp, Nkp = dims->Nkp, nG = dims->nG;
    if (Nip < 1 || Njp < 1 || Nkp < 1 || nG < 1 || nG > 16)
        return gam_fail_(GAM_E_ARG, "gam_divb_noalloc: bad sizes", "");
    const long nface = (Nip + 2 * nG + 1) * (Njp + 2 * nG + 1) * (Nkp + 2 * nG + 1);
    if (Nip * Njp * Nkp > GAM_DIVB_NOALLOC_MAX_CELLS || 3 * nface > (1L << 40))
        return gam_fail_(GAM_E_ARG, "gam_divb_noalloc: more active cells than the supported maximum", "");
    double *part = nullptr;
    int rc = scratch_of_current_device(&part);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_divb_noalloc, dim3(NA_BLOCKS), dim3(NA_THREADS), 0, s, (int)Nip, (int)Njp, (int)Nkp,
                       (int)nG, magFlux, bFlux0, part);
    hipLaunchKernelGGL(k_sum_noalloc, dim3(1), dim3(NA_THREADS), 0, s, (const double *)part, totDivB);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return gam_fail_(GAM_E_HIP, "gam_divb_noalloc: launch failed", hipGetErrorString(e));
    return GAM_OK;
}
