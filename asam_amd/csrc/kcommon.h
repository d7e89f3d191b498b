// kcommon.h -- device helpers shared by the stencil, line and set-up kernels.
#pragma once
#include <type_traits>

#include "fin.h"
#include "internal.h"

namespace asam {

// ===================================================================== helpers
template <int NQ>
__device__ __forceinline__ void block_reduce_store(double (&v)[NQ], double* partial, int blk,
                                                   int tid, int nthreads,
                                                   long stride = kRedBlocks) {
  __shared__ double sh[NQ][16];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
  }
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
