// blas.hip -- fused vector updates, deterministic two-stage reductions and the
// single-thread Krylov control tails (scalars never leave the device).
//
// Reduction stage 1 lives in the producing kernel (wavefront butterfly via
// __shfl_xor, then LDS across waves, one partial per workgroup); stage 2 is a
// one-workgroup finalize kernel that sums the partials in a fixed order, so a
// dot product is bitwise reproducible run to run.
#include <math.h>

#include "internal.h"

namespace asam {

template <int NQ>
__device__ __forceinline__ void block_reduce_store_b(double (&v)[NQ], double* partial, int blk) {
  __shared__ double sh[NQ][16];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
  }
  if ((tid & 63) == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) sh[q][tid >> 6] = v[q];
  }
  __syncthreads();
  if (tid == 0) {
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += sh[q][w];
      partial[q * kRedBlocks + blk] = s;
    }
  }
}

int blas_blocks(long n) {
  long b = (n + 1023) / 1024;
  if (b > 1024) b = 1024;
  if (b < 1) b = 1;
  return (int)b;
}

// ------------------------------------------------------------- BiCGStab tail
template <typename VT>
__global__ __launch_bounds__(256) void k_bicg_update(long n, VT* __restrict__ x,
                                                     const VT* __restrict__ ph,
                                                     const VT* __restrict__ rh,
                                                     VT* __restrict__ r, const VT* __restrict__ t,
                                                     const VT* __restrict__ rt, const KState* st,
                                                     double* partial) {
  if (st->done) return;
  const double alpha = st->alpha, omega = st->omega;
  const bool half = st->half != 0;
  double acc[2] = {0.0, 0.0};
  const long stride = (long)gridDim.x * blockDim.x;
  if (half) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride)
      x[e] = (VT)((double)x[e] + alpha * (double)ph[e]);
  } else {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride) {
      double xv = (double)x[e];
      if constexpr (sizeof(VT) == 8) {
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
  }
  block_reduce_store_b<2>(acc, partial, blockIdx.x);
}

void launch_bicg_update(hipStream_t s, long n, int vt, void* x, const void* ph, const void* rh,
                        void* r, const void* t, const void* rt, const KState* st,
                        double* partial) {
  int nb = blas_blocks(n);
  if (vt == 0)
    hipLaunchKernelGGL(k_bicg_update<double>, dim3(nb), dim3(256), 0, s, n, (double*)x,
                       (const double*)ph, (const double*)rh, (double*)r, (const double*)t,
                       (const double*)rt, st, partial);
  else
    hipLaunchKernelGGL(k_bicg_update<float>, dim3(nb), dim3(256), 0, s, n, (float*)x,
                       (const float*)ph, (const float*)rh, (float*)r, (const float*)t,
                       (const float*)rt, st, partial);
}

// ------------------------------------------------------------- (F)GMRES MGS
template <typename VT>
__global__ __launch_bounds__(256) void k_mgs(long n, VT* __restrict__ w,
                                             const VT* __restrict__ vprev,
                                             const VT* __restrict__ vk, int k, int i,
                                             const KState* st, double* partial, int chk_done) {
  if (chk_done ? st->done : st->cycle_done) return;
  double h = 0.0, sp = 0.0, sk = 0.0;
  if (k > 0) { h = st->H[(k - 1) * kMaxM + i]; sp = st->vscale[k - 1]; }
  const bool self = (k == i + 1);
  if (!self) sk = st->vscale[k];
  double acc[1] = {0.0};
  const long stride = (long)gridDim.x * blockDim.x;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride) {
    double wv = (double)w[e];
    if (k > 0) {
      wv = wv + (-h) * ((double)vprev[e] * sp);
      w[e] = (VT)wv;
      wv = (double)(VT)wv;
    }
    if (self) acc[0] += wv * wv;
    else acc[0] += ((double)vk[e] * sk) * wv;
  }
  block_reduce_store_b<1>(acc, partial, blockIdx.x);
}

void launch_mgs(hipStream_t s, long n, int vt, void* w, const void* vprev, const void* vk, int k,
                int i, const KState* st, double* partial, int chk_done) {
  int nb = blas_blocks(n);
  if (vt == 0)
    hipLaunchKernelGGL(k_mgs<double>, dim3(nb), dim3(256), 0, s, n, (double*)w,
                       (const double*)vprev, (const double*)vk, k, i, st, partial, chk_done);
  else
    hipLaunchKernelGGL(k_mgs<float>, dim3(nb), dim3(256), 0, s, n, (float*)w,
                       (const float*)vprev, (const float*)vk, k, i, st, partial, chk_done);
}

struct ZList {
  const void* z[kMaxM];
};

template <typename VT>
__global__ __launch_bounds__(256) void k_gmres_update(long n, VT* __restrict__ x, ZList zl,
                                                      int m, int use_vscale, const KState* st) {
  if (st->pad) return;  // pad doubles as "no update this cycle"
  const int nR = st->nR;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride) {
    double xv = (double)x[e];
    for (int q = 0; q < nR && q < m; ++q) {
      double zq = (double)((const VT*)zl.z[q])[e];
      if (use_vscale) zq = zq * st->vscale[q];
      xv = xv + st->y[q] * zq;
    }
    x[e] = (VT)xv;
  }
}

void launch_gmres_update(hipStream_t s, long n, int vt, void* x, void* const* zs, int m,
                         int use_vscale, const KState* st) {
  ZList zl;
  for (int q = 0; q < kMaxM; ++q) zl.z[q] = q < m ? zs[q] : nullptr;
  int nb = blas_blocks(n);
  if (vt == 0)
    hipLaunchKernelGGL(k_gmres_update<double>, dim3(nb), dim3(256), 0, s, n, (double*)x, zl, m,
                       use_vscale, st);
  else
    hipLaunchKernelGGL(k_gmres_update<float>, dim3(nb), dim3(256), 0, s, n, (float*)x, zl, m,
                       use_vscale, st);
}

// ------------------------------------------------------------- generic BLAS-1
__global__ __launch_bounds__(256) void k_dot(long n, const double* x, const double* y,
                                             double* partial) {
  double acc[1] = {0.0};
  const long stride = (long)gridDim.x * blockDim.x;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride)
    acc[0] += x[e] * y[e];
  block_reduce_store_b<1>(acc, partial, blockIdx.x);
}

void launch_dot(hipStream_t s, long n, const double* x, const double* y, double* partial) {
  hipLaunchKernelGGL(k_dot, dim3(blas_blocks(n)), dim3(256), 0, s, n, x, y, partial);
}

__global__ void k_axpby(long n, double a, const double* x, double b, double* y) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride)
    y[e] = a * x[e] + b * y[e];
}

void launch_axpby(hipStream_t s, long n, double a, const double* x, double b, double* y) {
  hipLaunchKernelGGL(k_axpby, dim3(blas_blocks(n)), dim3(256), 0, s, n, a, x, b, y);
}

template <typename VT>
__global__ void k_fill(long n, VT* x, double v) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride) x[e] = (VT)v;
}

void launch_fill(hipStream_t s, long n, int vt, void* x, double v) {
  if (vt == 0) hipLaunchKernelGGL(k_fill<double>, dim3(blas_blocks(n)), dim3(256), 0, s, n,
                                  (double*)x, v);
  else hipLaunchKernelGGL(k_fill<float>, dim3(blas_blocks(n)), dim3(256), 0, s, n, (float*)x,
                          v);
}

template <typename TD, typename TS>
__global__ void k_copy_cvt(long n, TD* d, const TS* s) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride)
    d[e] = (TD)s[e];
}

void launch_copy_cvt(hipStream_t s, long n, int vt_dst, void* dst, int vt_src, const void* src) {
  dim3 g(blas_blocks(n)), b(256);
  if (vt_dst == 0 && vt_src == 0)
    hipLaunchKernelGGL((k_copy_cvt<double, double>), g, b, 0, s, n, (double*)dst,
                       (const double*)src);
  else if (vt_dst == 0)
    hipLaunchKernelGGL((k_copy_cvt<double, float>), g, b, 0, s, n, (double*)dst,
                       (const float*)src);
  else if (vt_src == 0)
    hipLaunchKernelGGL((k_copy_cvt<float, double>), g, b, 0, s, n, (float*)dst,
                       (const double*)src);
  else
    hipLaunchKernelGGL((k_copy_cvt<float, float>), g, b, 0, s, n, (float*)dst,
                       (const float*)src);
}

__global__ void k_mass(long N, int np1, const double* vol, double* y, const double* x) {
  long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= N * np1) return;
  y[e] = vol[e / np1] * x[e];
}

void launch_mass(hipStream_t s, const Geom& g, const double* vol, double* y, const double* x) {
  long n = g.N * g.np1;
  hipLaunchKernelGGL(k_mass, dim3((n + 255) / 256), dim3(256), 0, s, g.N, g.np1, vol, y, x);
}

// ------------------------------------------------------------- finalize
__global__ __launch_bounds__(256) void k_finalize(int kind, int nq, int nblk,
                                                  const double* partial, KState* st, int k,
                                                  int i, double* out) {
  __shared__ double sh[4][256];
  const int tid = threadIdx.x;
  for (int q = 0; q < nq; ++q) {
    double s = 0.0;
    for (int b = tid; b < nblk; b += 256) s += partial[q * kRedBlocks + b];
    sh[q][tid] = s;
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
      for (int q = 0; q < nq; ++q) sh[q][tid] += sh[q][tid + w];
    __syncthreads();
  }
  if (tid != 0) return;
  double S[4];
  for (int q = 0; q < 4; ++q) S[q] = q < nq ? sh[q][0] : 0.0;
  switch (kind) {
    case FIN_SUMS:
      for (int q = 0; q < 4; ++q) st->sum[q] = S[q];
      break;
    case FIN_DOT:
      out[0] = S[0];
      break;
    case FIN_BICG_INIT: {
      // rNorm0 = ||r||; iteration 1 computes rho = rTilde.r = ||r||^2
      st->rnorm0 = sqrt(S[0]);
      st->rnorm = st->rnorm0;
      st->rho = S[0];
      st->iter = 1;
      st->done = 0; st->half = 0; st->pad = 0;
      st->status = 1;
      if (st->rho == 0.0) {  // reference: MaxIter = Iter, Tol = 0
        st->result = 1; st->rnorm = 0.0; st->status = 2; st->done = 1; st->pad = 1;
      }
      break;
    }
    case FIN_BICG_ALPHA:
      if (st->done) break;
      st->alpha = st->rho / S[1];
      break;
    case FIN_BICG_HALF: {
      if (st->done) break;
      double rn = sqrt(S[0]);
      st->rnorm = rn;
      if (rn <= st->tol * st->rnorm0) {
        st->half = 1; st->pad = 1; st->result = -st->iter; st->status = 0;
      }
      break;
    }
    case FIN_BICG_OMEGA:
      if (st->done || st->pad) break;
      st->omega = S[1] / S[0];
      break;
    case FIN_BICG_END: {
      if (st->done) break;
      if (st->half) { st->done = 1; st->pad = 1; break; }
      double rn = sqrt(S[0]);
      st->rnorm = rn;
      if (rn <= st->tol * st->rnorm0) {
        st->result = st->iter; st->status = 0; st->done = 1;
      } else if (st->iter == st->maxiter) {
        st->result = 2 * st->iter; st->status = 1; st->done = 1;
      } else {
        st->rho_old = st->rho;
        st->rho = S[1];
        if (st->rho == 0.0) {
          st->result = st->iter + 1; st->rnorm = 0.0; st->status = 2; st->done = 1;
        } else {
          st->beta = (st->rho / st->rho_old) * (st->alpha / st->omega);
          st->iter += 1;
        }
      }
      st->pad = st->done;
      break;
    }
    case FIN_GM_INIT: {
      // start of a restart cycle: beta = ||b - Op x||
      if (st->done) { st->pad = 1; st->cycle_done = 1; break; }
      double beta = sqrt(S[0]);
      if (st->cycles == 0) st->grho0 = beta;
      st->gbeta = beta;
      if (beta <= 0.0) {
        st->done = 1; st->pad = 1; st->cycle_done = 1; st->grho = 0.0; st->gstatus = 0;
        break;
      }
      st->vscale[0] = 1.0 / beta;
      st->bhat[0] = beta * 1.0;
      st->nR = st->m;
      st->cycle_done = 0;
      st->pad = 0;
      break;
    }
    case FIN_GM_H:
      if (st->cycle_done) break;
      st->H[k * kMaxM + i] = (nq == 2) ? S[1] * st->vscale[0] : S[0];
      break;
    case FIN_GM_LAST: {
      if (st->cycle_done) break;
      double* H = st->H;
      const int M = kMaxM;
      H[(i + 1) * M + i] = sqrt(S[0]);
      st->vscale[i + 1] = 1.0 / H[(i + 1) * M + i];
      for (int q = 1; q <= i; ++q) {
        double temp = st->cs[q - 1] * H[(q - 1) * M + i] + st->sn[q - 1] * H[q * M + i];
        H[q * M + i] = -st->sn[q - 1] * H[(q - 1) * M + i] + st->cs[q - 1] * H[q * M + i];
        H[(q - 1) * M + i] = temp;
      }
      double delta = sqrt(H[i * M + i] * H[i * M + i] + H[(i + 1) * M + i] * H[(i + 1) * M + i]);
      st->cs[i] = H[i * M + i] / delta;
      st->sn[i] = H[(i + 1) * M + i] / delta;
      H[i * M + i] = st->cs[i] * H[i * M + i] + st->sn[i] * H[(i + 1) * M + i];
      st->bhat[i + 1] = -st->sn[i] * st->bhat[i];
      st->bhat[i] = st->cs[i] * st->bhat[i];
      st->grho = fabs(st->bhat[i + 1]);
      st->total += 1;
      if (st->grho <= st->gtol * st->grho0) {
        st->nR = i + 1;
        st->cycle_done = 1;
      } else {
        st->nR = st->m;
        if (i == st->m - 1) st->cycle_done = 1;
      }
      break;
    }
    default:
      break;
  }
}

void launch_finalize(hipStream_t s, int kind, int nq, int nblk, const double* partial,
                     KState* st, int k, int i, double* out) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, kind, nq, nblk, partial, st, k, i,
                     out);
}

// (F)GMRES back-substitution (Iter_Mod.f90:181-188) and end-of-cycle decision.
__global__ void k_gm_backsolve(KState* st) {
  if (threadIdx.x != 0) return;
  if (st->pad) return;  // no cycle ran
  const int nR = st->nR, M = kMaxM;
  const double* H = st->H;
  double* y = st->y;
  y[nR - 1] = st->bhat[nR - 1] / H[(nR - 1) * M + (nR - 1)];
  for (int k = nR - 2; k >= 0; --k) {
    y[k] = st->bhat[k];
    for (int q = k + 1; q < nR; ++q) y[k] = y[k] - H[k * M + q] * y[q];
    y[k] = y[k] / H[k * M + k];
  }
  st->cycles += 1;
  if (st->grho <= st->gtol * st->grho0) {
    st->done = 1; st->gstatus = 0;
  } else if (st->cycles >= st->maxcycles) {
    st->done = 1; st->gstatus = 1;
  }
}

void launch_gm_backsolve(hipStream_t s, KState* st) {
  hipLaunchKernelGGL(k_gm_backsolve, dim3(1), dim3(64), 0, s, st);
}

// Mixed refinement: residual norm of the FP64 outer residual in sum (partial q0).
__global__ __launch_bounds__(256) void k_mixed_outer(int nblk, const double* partial,
                                                     KState* st) {
  __shared__ double sh[256];
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int b = tid; b < nblk; b += 256) s += partial[b];
  sh[tid] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) sh[tid] += sh[tid + w];
    __syncthreads();
  }
  if (tid != 0) return;
  if (st->outer_done) return;
  double r = sqrt(sh[0]);
  if (st->outer == 0) st->outer_r0 = r;
  st->outer_r = r;
  if (r <= st->outer_tol * st->outer_r0 || r == 0.0) {
    st->outer_done = 1;
    st->gstatus = 0;
    st->done = 1; st->pad = 1;
    return;
  }
  // Inner BiCGStab on A d = r (d = 0): rNorm0 = ||r||, rho = ||r||^2 (rTilde = r).
  // Ask for the inner inner_tol reduction, but not more than the outer target needs.
  double need = 0.5 * st->outer_tol * st->outer_r0 / r;
  double itol = st->inner_tol;
  if (need > itol) itol = need;
  if (itol > 0.5) itol = 0.5;
  st->tol = itol;
  st->rnorm0 = r;
  st->rnorm = r;
  st->rho = sh[0];
  st->iter = 1;
  st->done = 0; st->half = 0; st->pad = 0; st->status = 1;
  st->outer += 1;
}

void launch_mixed_outer(hipStream_t s, int nblk, const double* partial, KState* st) {
  hipLaunchKernelGGL(k_mixed_outer, dim3(1), dim3(256), 0, s, nblk, partial, st);
}

}  // namespace asam
