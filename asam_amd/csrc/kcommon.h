// kcommon.h -- device helpers shared by the stencil, line and set-up kernels.
#pragma once
#include <type_traits>

#include "fin.h"
#include "internal.h"

namespace asam {

// ===================================================================== helpers
// Sum over the 64 lanes of a wavefront in registers (DPP): swaps of lanes 1 and 2 apart,
// the half-row mirror and the row mirror leave every lane of a 16-lane row holding the row
// total; the four row totals are then read out (readlane) and added in a fixed order, so
// the result is the same in every lane and bitwise reproducible.  All 64 lanes must be
// active at the call.
template <int CTRL>
__device__ __forceinline__ double dpp_add(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return v + __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double row_lane(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double wave_sum(double v) {
  v = dpp_add<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
  v = dpp_add<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
  v = dpp_add<0x141>(v);  // row_half_mirror: the two quads of each 8-lane half
  v = dpp_add<0x140>(v);  // row_mirror: the two halves of each 16-lane row
  return (row_lane(v, 0) + row_lane(v, 16)) + (row_lane(v, 32) + row_lane(v, 48));
}

template <int NQ>
__device__ __forceinline__ void block_reduce_store(double (&v)[NQ], double* partial, int blk,
                                                   int tid, int nthreads,
                                                   long stride = kRedBlocks) {
  __shared__ double sh[NQ][16];
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] = wave_sum(v[q]);
  const int lane = tid & 63, wid = tid >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) sh[q][wid] = v[q];
  }
  __syncthreads();
  if (tid == 0) {
    const int nw = (nthreads + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += sh[q][w];
      partial[q * stride + blk] = s;
    }
  }
}

template <typename T> struct V2 { T a, b; };

template <typename T>
__device__ __forceinline__ V2<T> ld2(const T* p, long c) {
  if constexpr (sizeof(T) == 8) {
    double2 v = *reinterpret_cast<const double2*>(p + 2 * c);
    return {(T)v.x, (T)v.y};
  } else {
    float2 v = *reinterpret_cast<const float2*>(p + 2 * c);
    return {(T)v.x, (T)v.y};
  }
}
template <typename T, typename U>
__device__ __forceinline__ void st2(T* p, long c, U a, U b) {
  if constexpr (sizeof(T) == 8) {
    *reinterpret_cast<double2*>(p + 2 * c) = make_double2((double)a, (double)b);
  } else {
    *reinterpret_cast<float2*>(p + 2 * c) = make_float2((float)a, (float)b);
  }
}

// BiCGStab tail of one element (SpMatrix4CellPar_Mod.f90:322-331): x += alpha pHat +
// omega rHat; r = s - omega t; accumulates r.r and rTilde.r.  FP32 vectors update in
// FP32, the dots accumulate in FP64.
template <typename VT>
__device__ __forceinline__ void bicg_upd_elem(long e, VT* __restrict__ x,
                                              const VT* __restrict__ ph,
                                              const VT* __restrict__ rh, VT* __restrict__ r,
                                              const VT* __restrict__ t,
                                              const VT* __restrict__ rt, double alpha,
                                              double omega, double (&acc)[2]) {
  if constexpr (sizeof(VT) == 8) {
    double xv = (double)x[e];
    xv = xv + alpha * (double)ph[e];
    xv = xv + omega * (double)rh[e];
    double rv = (double)r[e] - omega * (double)t[e];
    x[e] = (VT)xv;
    r[e] = (VT)rv;
    acc[0] += rv * rv;
    acc[1] += (double)rt[e] * rv;
  } else {
    float xf = (float)x[e];
    xf = xf + (float)alpha * (float)ph[e];
    xf = xf + (float)omega * (float)rh[e];
    float rf = (float)r[e] - (float)omega * (float)t[e];
    x[e] = (VT)xf;
    r[e] = (VT)rf;
    acc[0] += (double)rf * (double)rf;
    acc[1] += (double)rt[e] * (double)rf;
  }
}

}  // namespace asam
