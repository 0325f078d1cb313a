// Host stand-in for <hip/hip_runtime.h>: lets the per-point device
// functions of mage_amd/csrc (gam_phys.h, gam_adj.h) compile with g++ for
// CPU tests.  Only what those headers use is defined.
#pragma once
#include <cmath>
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(x)
static inline double rsqrt(double x) { return 1.0 / std::sqrt(x); }
struct uint3 { unsigned x, y, z; };
