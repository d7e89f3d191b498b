// blas.hip -- fused vector updates, deterministic two-stage reductions and the
// single-thread Krylov control tails (scalars never leave the device).
//
// Reduction stage 1 lives in the producing kernel (wavefront sum in registers via DPP,
// then LDS across waves, one partial per workgroup); stage 2 is a
// one-workgroup finalize kernel that sums the partials in a fixed order, so a
// dot product is bitwise reproducible run to run.
#include <math.h>
#include <stdint.h>

#include "kcommon.h"

namespace asam {

template <int NQ>
__device__ __forceinline__ void block_reduce_store_b(double (&v)[NQ], double* partial, int blk) {
  __shared__ double sh[NQ][16];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] = wave_sum(v[q]);
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
      // written through to the device-coherent level, for a fused tail (kcommon.h)
      __hip_atomic_store(partial + q * kRedBlocks + blk, s, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
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
                                                     double* partial, KTail tail, KTail fold) {
  if (st->done) return;  // the state before this launch: the same in every workgroup
  // a folded tail (the half-step test and omega): the scalars come from this workgroup's
  // copy of the state
  __shared__ unsigned long long sk[kTailWords];
  if (fold.kind >= 0) {
    fold_state_256(fold, sk, threadIdx.x);
    fold_write(fold, sk, threadIdx.x, 256);
    st = reinterpret_cast<const KState*>(sk);
  }
  const double alpha = st->alpha, omega = st->omega;
  const bool half = st->half != 0;
  double acc[2] = {0.0, 0.0};
  const long stride = (long)gridDim.x * blockDim.x;
  if (half) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride)
      x[e] = (VT)((double)x[e] + alpha * (double)ph[e]);
  } else {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride)
      bicg_upd_elem<VT>(e, x, ph, rh, r, t, rt, alpha, omega, acc);
  }
  block_reduce_store_b<2>(acc, partial, blockIdx.x);
  if (tail.kind >= 0) tail_arrive(tail, threadIdx.x);
}

void launch_bicg_update(hipStream_t s, long n, int vt, void* x, const void* ph, const void* rh,
                        void* r, const void* t, const void* rt, const KState* st,
                        double* partial, const KTail& tail, const KTail& fold) {
  int nb = blas_blocks(n);
  if (vt == 0)
    hipLaunchKernelGGL(k_bicg_update<double>, dim3(nb), dim3(256), 0, s, n, (double*)x,
                       (const double*)ph, (const double*)rh, (double*)r, (const double*)t,
                       (const double*)rt, st, partial, tail, fold);
  else
    hipLaunchKernelGGL(k_bicg_update<float>, dim3(nb), dim3(256), 0, s, n, (float*)x,
                       (const float*)ph, (const float*)rh, (float*)r, (const float*)t,
                       (const float*)rt, st, partial, tail, fold);
}

// ------------------------------------------------------------- (F)GMRES MGS
// Sum of the nb partial products of the previous pass, in the same fixed order in every
// workgroup (so every workgroup holds the same bits); all threads must take part.
__device__ double block_sum_partials(const double* p, int nb) {
  __shared__ double sh[4];
  __shared__ double total;
  const int tid = threadIdx.x;
  double v = 0.0;
  for (int b = tid; b < nb; b += 256) v += p[b];
  v = wave_sum(v);
  if ((tid & 63) == 0) sh[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) total = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return total;
}

// 16-byte packs of a vector: W = 4 floats or 2 doubles per load and store.  The FP32
// passes of the Gram-Schmidt chains are latency-bound with one element per lane (half the
// bytes in flight of the FP64 passes at the same speed), so the kernels below stream
// packs when every vector they touch is 16-byte aligned.
template <typename VT>
struct Pk {
  static constexpr int W = 16 / (int)sizeof(VT);
  VT v[W];
};
template <typename VT>
__device__ __forceinline__ Pk<VT> ldpk(const VT* p, long c) {
  Pk<VT> r;
  if constexpr (sizeof(VT) == 4) {
    const float4 t = reinterpret_cast<const float4*>(p)[c];
    r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z; r.v[3] = t.w;
  } else {
    const double2 t = reinterpret_cast<const double2*>(p)[c];
    r.v[0] = t.x; r.v[1] = t.y;
  }
  return r;
}
template <typename VT>
__device__ __forceinline__ void stpk(VT* p, long c, const Pk<VT>& r) {
  if constexpr (sizeof(VT) == 4)
    reinterpret_cast<float4*>(p)[c] = make_float4(r.v[0], r.v[1], r.v[2], r.v[3]);
  else
    reinterpret_cast<double2*>(p)[c] = make_double2(r.v[0], r.v[1]);
}
template <typename VT>
__device__ __forceinline__ Pk<VT> zpk() {
  Pk<VT> r;
#pragma unroll
  for (int j = 0; j < Pk<VT>::W; ++j) r.v[j] = (VT)0;
  return r;
}
__host__ inline bool al16(const void* p) { return ((uintptr_t)p & 15) == 0; }

// One element of a modified Gram-Schmidt pass: w -= h (v_prev * sp) (k > 0), rounded to
// the vector type; then the product with v_k * sk, or w.w on the step's last pass.
template <typename VT>
__device__ __forceinline__ VT mgs_elem(VT w, VT vp, VT vk, int k, bool self, double h,
                                       double sp, double sk, double& acc) {
  double wv = (double)w;
  if (k > 0) {
    wv = wv + (-h) * ((double)vp * sp);
    w = (VT)wv;
    wv = (double)w;
  }
  if (self) acc += wv * wv;
  else acc += ((double)vk * sk) * wv;
  return w;
}

// hsrc: null, or the partial products of the previous pass of the chain; then H(k-1, i) is
// their sum, formed here (workgroup 0 records it) instead of by a finalize launch.
// PK: the vectors stream in 16-byte packs (every pointer aligned), the n mod W last
// elements one per thread.
template <typename VT, bool PK>
__global__ __launch_bounds__(256) void k_mgs(long n, VT* __restrict__ w,
                                             const VT* __restrict__ vprev,
                                             const VT* __restrict__ vk, int k, int i,
                                             KState* st, double* partial, int chk_done,
                                             const double* hsrc, int hnb) {
  if (chk_done ? st->done : st->cycle_done) return;
  double h = 0.0, sp = 0.0, sk = 0.0;
  if (k > 0) {
    if (hsrc) {
      h = block_sum_partials(hsrc, hnb);
      if (blockIdx.x == 0 && threadIdx.x == 0) st->H[(k - 1) * kMaxM + i] = h;
    } else {
      h = st->H[(k - 1) * kMaxM + i];
    }
    sp = st->vscale[k - 1];
  }
  const bool self = (k == i + 1);
  if (!self) sk = st->vscale[k];
  double acc[1] = {0.0};
  const long stride = (long)gridDim.x * blockDim.x;
  const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if constexpr (PK) {
    constexpr int W = Pk<VT>::W, U = 2;
    const long nc = n / W;
    long c0 = gid;
    // whole turns: U packs per thread, all loads first
    for (; c0 + (U - 1) * stride < nc; c0 += U * stride) {
      Pk<VT> wl[U], pl[U], kl[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long c = c0 + u * stride;
        wl[u] = ldpk(w, c);
        pl[u] = k > 0 ? ldpk(vprev, c) : zpk<VT>();
        kl[u] = !self ? ldpk(vk, c) : zpk<VT>();
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
#pragma unroll
        for (int j = 0; j < W; ++j)
          wl[u].v[j] = mgs_elem(wl[u].v[j], pl[u].v[j], kl[u].v[j], k, self, h, sp, sk, acc[0]);
        if (k > 0) stpk(w, c0 + u * stride, wl[u]);
      }
    }
    for (; c0 < nc; c0 += stride) {
      Pk<VT> wl = ldpk(w, c0);
      const Pk<VT> pl = k > 0 ? ldpk(vprev, c0) : zpk<VT>();
      const Pk<VT> kl = !self ? ldpk(vk, c0) : zpk<VT>();
#pragma unroll
      for (int j = 0; j < W; ++j)
        wl.v[j] = mgs_elem(wl.v[j], pl.v[j], kl.v[j], k, self, h, sp, sk, acc[0]);
      if (k > 0) stpk(w, c0, wl);
    }
    const long e = nc * W + gid;
    if (e < n) {
      const VT r = mgs_elem(w[e], k > 0 ? vprev[e] : (VT)0, !self ? vk[e] : (VT)0, k, self, h,
                            sp, sk, acc[0]);
      if (k > 0) w[e] = r;
    }
  } else {
    // U turns of the grid-stride loop at a time: the loads of all U elements first, then
    // their arithmetic in the plain loop's order (same elements, same sums, same bits)
    constexpr int U = 4;
    for (long e0 = gid; e0 < n; e0 += U * stride) {
      VT wl[U], pl[U], kl[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long e = e0 + u * stride;
        const bool ok = e < n;
        wl[u] = ok ? w[e] : (VT)0;
        pl[u] = (ok && k > 0) ? vprev[e] : (VT)0;
        kl[u] = (ok && !self) ? vk[e] : (VT)0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long e = e0 + u * stride;
        if (e < n) {
          const VT r = mgs_elem(wl[u], pl[u], kl[u], k, self, h, sp, sk, acc[0]);
          if (k > 0) w[e] = r;
        }
      }
    }
  }
  block_reduce_store_b<1>(acc, partial, blockIdx.x);
}

void launch_mgs(hipStream_t s, long n, int vt, void* w, const void* vprev, const void* vk, int k,
                int i, KState* st, double* partial, int chk_done, const double* hsrc) {
  const int nb = blas_blocks(n);
  const bool pk = al16(w) && (k == 0 || al16(vprev)) && (!vk || al16(vk));
  if (vt == 0) {
    auto f = pk ? k_mgs<double, true> : k_mgs<double, false>;
    hipLaunchKernelGGL(f, dim3(nb), dim3(256), 0, s, n, (double*)w, (const double*)vprev,
                       (const double*)vk, k, i, st, partial, chk_done, hsrc, nb);
  } else {
    auto f = pk ? k_mgs<float, true> : k_mgs<float, false>;
    hipLaunchKernelGGL(f, dim3(nb), dim3(256), 0, s, n, (float*)w, (const float*)vprev,
                       (const float*)vk, k, i, st, partial, chk_done, hsrc, nb);
  }
}

struct ZList {
  const void* z[kMaxM];
};

// x += sum_q y_q z_q (times vscale_q for GMRES's unscaled basis), one element.
template <typename VT>
__device__ __forceinline__ double gm_upd_elem(double xv, double zq, int use_vscale,
                                              double vsq, double yq) {
  if (use_vscale) zq = zq * vsq;
  return xv + yq * zq;
}

template <typename VT, bool PK>
__global__ __launch_bounds__(256) void k_gmres_update(long n, VT* __restrict__ x, ZList zl,
                                                      int m, int use_vscale, const KState* st) {
  if (st->pad) return;  // pad doubles as "no update this cycle"
  const int nR = st->nR;
  const long stride = (long)gridDim.x * blockDim.x;
  const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if constexpr (PK) {
    constexpr int W = Pk<VT>::W;
    const long nc = n / W;
    for (long c = gid; c < nc; c += stride) {
      Pk<VT> xl = ldpk(x, c);
      double xv[W];
#pragma unroll
      for (int j = 0; j < W; ++j) xv[j] = (double)xl.v[j];
      for (int q = 0; q < nR && q < m; ++q) {
        const Pk<VT> zp = ldpk((const VT*)zl.z[q], c);
#pragma unroll
        for (int j = 0; j < W; ++j)
          xv[j] = gm_upd_elem<VT>(xv[j], (double)zp.v[j], use_vscale, st->vscale[q], st->y[q]);
      }
#pragma unroll
      for (int j = 0; j < W; ++j) xl.v[j] = (VT)xv[j];
      stpk(x, c, xl);
    }
    const long e = nc * W + gid;
    if (e < n) {
      double xv = (double)x[e];
      for (int q = 0; q < nR && q < m; ++q)
        xv = gm_upd_elem<VT>(xv, (double)((const VT*)zl.z[q])[e], use_vscale, st->vscale[q],
                             st->y[q]);
      x[e] = (VT)xv;
    }
  } else {
    for (long e = gid; e < n; e += stride) {
      double xv = (double)x[e];
      for (int q = 0; q < nR && q < m; ++q)
        xv = gm_upd_elem<VT>(xv, (double)((const VT*)zl.z[q])[e], use_vscale, st->vscale[q],
                             st->y[q]);
      x[e] = (VT)xv;
    }
  }
}

void launch_gmres_update(hipStream_t s, long n, int vt, void* x, void* const* zs, int m,
                         int use_vscale, const KState* st) {
  ZList zl;
  bool pk = al16(x);
  for (int q = 0; q < kMaxM; ++q) {
    zl.z[q] = q < m ? zs[q] : nullptr;
    if (q < m) pk = pk && al16(zs[q]);
  }
  int nb = blas_blocks(n);
  if (vt == 0) {
    auto f = pk ? k_gmres_update<double, true> : k_gmres_update<double, false>;
    hipLaunchKernelGGL(f, dim3(nb), dim3(256), 0, s, n, (double*)x, zl, m, use_vscale, st);
  } else {
    auto f = pk ? k_gmres_update<float, true> : k_gmres_update<float, false>;
    hipLaunchKernelGGL(f, dim3(nb), dim3(256), 0, s, n, (float*)x, zl, m, use_vscale, st);
  }
}

// ------------------------------------------------------- FGMRESV classical Gram-Schmidt
// FGMRESV (Iter_Mod.f90:201-288) takes every projection H(k,i) = v_k . w of an Arnoldi
// step from the same w (:238-241) before removing any (:242-244), so the projections need
// no pass between them: up to 4 come from one read of w.  The basis is stored unscaled,
// v_k = vs[k] * vscale[k], as in the modified Gram-Schmidt path.
template <typename VT, bool PK>
__global__ __launch_bounds__(256) void k_cgs_dots(long n, const VT* __restrict__ w, ZList vl,
                                                  int k0, int nq, const KState* st,
                                                  double* partial) {
  if (st->cycle_done) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0}, sk[4];
  const VT* v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    sk[q] = q < nq ? st->vscale[k0 + q] : 0.0;
    v[q] = q < nq ? (const VT*)vl.z[k0 + q] : w;
  }
  const long stride = (long)gridDim.x * blockDim.x;
  const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if constexpr (PK) {
    constexpr int W = Pk<VT>::W;
    const long nc = n / W;
    for (long c = gid; c < nc; c += stride) {
      const Pk<VT> wp = ldpk(w, c);
      Pk<VT> vp[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) vp[q] = q < nq ? ldpk(v[q], c) : zpk<VT>();
#pragma unroll
      for (int j = 0; j < W; ++j) {
        const double wv = (double)wp.v[j];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (q < nq) acc[q] += ((double)vp[q].v[j] * sk[q]) * wv;
      }
    }
    const long e = nc * W + gid;
    if (e < n) {
      const double wv = (double)w[e];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (q < nq) acc[q] += ((double)v[q][e] * sk[q]) * wv;
    }
  } else {
    constexpr int U = 4;  // U grid-stride turns at a time, loads first (same order and sums)
    for (long e0 = gid; e0 < n; e0 += U * stride) {
      VT wl[U], vl4[4][U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long e = e0 + u * stride;
        const bool ok = e < n;
        wl[u] = ok ? w[e] : (VT)0;
#pragma unroll
        for (int q = 0; q < 4; ++q) vl4[q][u] = (ok && q < nq) ? v[q][e] : (VT)0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (e0 + u * stride < n) {
          const double wv = (double)wl[u];
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (q < nq) acc[q] += ((double)vl4[q][u] * sk[q]) * wv;
        }
      }
    }
  }
  block_reduce_store_b<4>(acc, partial, blockIdx.x);
}

// w -= H(k,i) v_k for k = 0..i in that order (the DAXPYs of Iter_Mod.f90:242-244, each
// rounded to the vector type as a separate update), then |w|^2.
template <typename VT, bool PK>
__global__ __launch_bounds__(256) void k_cgs_update(long n, VT* __restrict__ w, ZList vl,
                                                    int i, const KState* st, double* partial) {
  if (st->cycle_done) return;
  double mh[kMaxM], sk[kMaxM];
#pragma unroll
  for (int k = 0; k < kMaxM; ++k) {
    mh[k] = k <= i ? -st->H[k * kMaxM + i] : 0.0;
    sk[k] = k <= i ? st->vscale[k] : 0.0;
  }
  double acc[1] = {0.0};
  const long stride = (long)gridDim.x * blockDim.x;
  const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if constexpr (PK) {
    constexpr int W = Pk<VT>::W;
    const long nc = n / W;
    for (long c = gid; c < nc; c += stride) {
      Pk<VT> wp = ldpk(w, c);
      double wv[W];
#pragma unroll
      for (int j = 0; j < W; ++j) wv[j] = (double)wp.v[j];
#pragma unroll
      for (int k = 0; k < kMaxM; ++k)
        if (k <= i) {
          const Pk<VT> vp = ldpk((const VT*)vl.z[k], c);
#pragma unroll
          for (int j = 0; j < W; ++j)
            wv[j] = (double)(VT)(wv[j] + mh[k] * ((double)vp.v[j] * sk[k]));
        }
#pragma unroll
      for (int j = 0; j < W; ++j) {
        wp.v[j] = (VT)wv[j];
        acc[0] += wv[j] * wv[j];
      }
      stpk(w, c, wp);
    }
    const long e = nc * W + gid;
    if (e < n) {
      double wv = (double)w[e];
#pragma unroll
      for (int k = 0; k < kMaxM; ++k)
        if (k <= i) wv = (double)(VT)(wv + mh[k] * ((double)((const VT*)vl.z[k])[e] * sk[k]));
      w[e] = (VT)wv;
      acc[0] += wv * wv;
    }
  } else {
    for (long e = gid; e < n; e += stride) {
      double wv = (double)w[e];
#pragma unroll
      for (int k = 0; k < kMaxM; ++k)
        if (k <= i) wv = (double)(VT)(wv + mh[k] * ((double)((const VT*)vl.z[k])[e] * sk[k]));
      w[e] = (VT)wv;
      acc[0] += wv * wv;
    }
  }
  block_reduce_store_b<1>(acc, partial, blockIdx.x);
}

void launch_cgs_dots(hipStream_t s, long n, int vt, const void* w, void* const* vs, int k0,
                     int nq, const KState* st, double* partial) {
  ZList vl;  // vs holds at least k0 + nq vectors
  bool pk = al16(w);
  for (int q = 0; q < kMaxM; ++q) {
    vl.z[q] = q < k0 + nq ? vs[q] : nullptr;
    if (q >= k0 && q < k0 + nq) pk = pk && al16(vs[q]);
  }
  if (vt == 0) {
    auto f = pk ? k_cgs_dots<double, true> : k_cgs_dots<double, false>;
    hipLaunchKernelGGL(f, dim3(blas_blocks(n)), dim3(256), 0, s, n, (const double*)w, vl, k0,
                       nq, st, partial);
  } else {
    auto f = pk ? k_cgs_dots<float, true> : k_cgs_dots<float, false>;
    hipLaunchKernelGGL(f, dim3(blas_blocks(n)), dim3(256), 0, s, n, (const float*)w, vl, k0,
                       nq, st, partial);
  }
}

void launch_cgs_update(hipStream_t s, long n, int vt, void* w, void* const* vs, int i,
                       const KState* st, double* partial) {
  ZList vl;  // vs holds at least i + 1 vectors
  bool pk = al16(w);
  for (int q = 0; q < kMaxM; ++q) {
    vl.z[q] = q <= i ? vs[q] : nullptr;
    if (q <= i) pk = pk && al16(vs[q]);
  }
  if (vt == 0) {
    auto f = pk ? k_cgs_update<double, true> : k_cgs_update<double, false>;
    hipLaunchKernelGGL(f, dim3(blas_blocks(n)), dim3(256), 0, s, n, (double*)w, vl, i, st,
                       partial);
  } else {
    auto f = pk ? k_cgs_update<float, true> : k_cgs_update<float, false>;
    hipLaunchKernelGGL(f, dim3(blas_blocks(n)), dim3(256), 0, s, n, (float*)w, vl, i, st,
                       partial);
  }
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

// Any entry != 0 (a signed zero counts as zero): the lanes that see one store the flags;
// they start as xnz = 0, xz = 1 and are only ever set to the same values.
__global__ __launch_bounds__(256) void k_any_nonzero(long n, const double* __restrict__ x,
                                                     int* xnz, int* xz) {
  const long stride = (long)gridDim.x * blockDim.x;
  bool nz = false;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += stride)
    nz |= x[e] != 0.0;
  if (nz) {
    *xnz = 1;
    *xz = 0;
  }
}

void launch_any_nonzero(hipStream_t s, long n, const double* x, int* xnz, int* xz) {
  hipLaunchKernelGGL(k_any_nonzero, dim3(blas_blocks(n)), dim3(256), 0, s, n, x, xnz, xz);
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
// One workgroup of 1024 threads: every thread sums a fixed strided subset of the
// partials, then a fixed wavefront butterfly and a fixed 16-wave sum -> identical
// results run to run.  Thread 0 then runs the scalar step.
__global__ __launch_bounds__(1024) void k_finalize(int kind, int nq, int nblk, int nblk_hi,
                                                   const double* partial, KState* st, int k,
                                                   int i, double* out) {
  // The scalar step runs on an LDS copy of the solver state: one cooperative load and one
  // cooperative store instead of a chain of dependent global accesses by one thread (the
  // state was last written by the previous tail, usually from another XCD's L2).
  __shared__ KState sk;
  __shared__ double sh[4][16];
  static_assert(sizeof(KState) % 8 == 0, "KState is copied in 8-byte words");
  static_assert(offsetof(KState, H) % 8 == 0, "the scalar part is copied in 8-byte words");
  // the BiCGStab and refinement tails read and write only the scalars ahead of the
  // (F)GMRES arrays (about 350 bytes of the state's 10 KiB)
  const bool arrays = kind == FIN_GM_INIT || kind == FIN_GM_H || kind == FIN_GM_HV ||
                      kind == FIN_GM_LAST || kind == FIN_MIXED_GM;
  const int kWords = arrays ? (int)(sizeof(KState) / 8) : (int)(offsetof(KState, H) / 8);
  const int tid = threadIdx.x;
  for (int w = tid; w < kWords; w += 1024)
    reinterpret_cast<unsigned long long*>(&sk)[w] =
        reinterpret_cast<const unsigned long long*>(st)[w];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q < nq) {
      const int nb = q < 2 ? nblk : nblk_hi;
      const double* pq = partial + (long)q * kRedBlocks;
      double s = 0.0;
      for (int b = tid; b < nb; b += 1024) s += pq[b];
      v[q] = s;
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = wave_sum(v[q]);
  if ((tid & 63) == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) sh[q][tid >> 6] = v[q];
  }
  __syncthreads();
  if (tid == 0) {
    double S[4];
    for (int q = 0; q < 4; ++q) {
      double s = 0.0;
      for (int w = 0; w < 16; ++w) s += sh[q][w];
      S[q] = q < nq ? s : 0.0;
    }
    fin_logic(kind, S, &sk, k, i, nq, out);
  }
  __syncthreads();
  for (int w = tid; w < kWords; w += 1024)
    reinterpret_cast<unsigned long long*>(st)[w] =
        reinterpret_cast<const unsigned long long*>(&sk)[w];
}

void launch_finalize(hipStream_t s, int kind, int nq, int nblk, const double* partial,
                     KState* st, int k, int i, double* out, int nblk_hi) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, s, kind, nq, nblk,
                     nblk_hi > 0 ? nblk_hi : nblk, partial, st, k, i, out);
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

}  // namespace asam
