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
      // written through to the device-coherent level (a fused tail in another XCD may
      // read it before this kernel ends)
      __hip_atomic_store(partial + q * stride + blk, s, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Scalar words of the state ahead of the (F)GMRES arrays: all that the fused tails touch.
constexpr int kTailWords = (int)(offsetof(KState, H) / 8);

// The summed quantities S[q] of the partial sums t describes, in k_finalize's order, by a
// workgroup of 256 threads (see tail_arrive); slots outside t.qmask are 0.
__device__ __forceinline__ void tail_sums_256(const KTail& t, int tid, double (&S)[4]) {
  __shared__ double sh[4][16];
  __shared__ double vs[1024];
  const int lane = tid & 63, wid = tid >> 6;
  for (int q = 0; q < 4; ++q) {
    if (q >= t.nq || !((t.qmask >> q) & 1)) continue;  // uniform
    double s4[4] = {0.0, 0.0, 0.0, 0.0};
    const int nb = q < 2 ? t.nblk : t.nblk_hi;
    const double* pq = t.partial + (long)q * kRedBlocks;
    // full rounds of 1024 first (no lane condition, so their loads are in flight
    // together), then the partial last round
    const int full = nb >> 10;
#pragma unroll 4
    for (int r = 0; r < full; ++r) {
#pragma unroll
      for (int c = 0; c < 4; ++c) s4[c] += pq[(r << 10) + tid + (c << 8)];
    }
    if ((full << 10) < nb) {
      // clamped unconditional loads, the value selected after (adding +0.0 keeps the bits)
      double x[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) x[c] = pq[min((full << 10) + tid + (c << 8), nb - 1)];
#pragma unroll
      for (int c = 0; c < 4; ++c) s4[c] += ((full << 10) + tid + (c << 8) < nb) ? x[c] : 0.0;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) vs[tid + (c << 8)] = s4[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int g = wid + (c << 2);
      const double v = wave_sum(vs[(g << 6) + lane]);
      if (lane == 0) sh[q][g] = v;
    }
    __syncthreads();
  }
  for (int q = 0; q < 4; ++q) {
    double s = 0.0;
    if (q < t.nq && ((t.qmask >> q) & 1))
      for (int w = 0; w < 16; ++w) s += sh[q][w];
    S[q] = s;
  }
}

// Fused scalar tail, in place of a k_finalize launch.  Every thread of every workgroup
// of 256 threads calls it right after block_reduce_store.
//
// Publication: thread 0 has written the workgroup's partial sums with agent-scope
// (write-through) stores.  It waits for them to complete, then counts the workgroup in
// with an agent-scope atomic add.  This is the order an agent-scope release gives atomic
// stores; the release fence itself would also write back every dirty line of this XCD's
// L2 in every workgroup, which costs far more than the finalize launch it replaces.
//
// The workgroup that arrives last invalidates its caches (agent-scope acquire, after its
// atomic returned) and adds all partial sums in k_finalize's order: k_finalize's thread v
// (of 1024) sums the partials v, v + 1024, ... in turn; the butterfly then adds each group
// of 64 such sums; the 16 group totals are added in order.  Here thread t forms the sums
// of v = t, t + 256, t + 512, t + 768 side by side (independent loads in flight), parks
// them in LDS, and each wave runs the butterflies of groups w, w + 4, w + 8, w + 12, so
// the scalars have k_finalize's bits.  It runs the scalar step on an LDS copy of the
// state's scalars and resets the counter.  No workgroup waits for another.
__device__ __forceinline__ void tail_arrive(const KTail& t, int tid) {
  __shared__ int s_last;
  __shared__ unsigned long long sk[kTailWords];
  if (tid == 0) {
    // completion of this wave's stores, the partial sums included; also a compiler
    // barrier, so the atomic add below is not issued before them
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    const unsigned int total = gridDim.x * gridDim.y * gridDim.z;
    const int last = __hip_atomic_fetch_add(t.counter, 1u, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT) == total - 1u ? 1 : 0;
    // the barrier below orders the loads of the workgroup's other threads after this
    if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    s_last = last;
  }
  __syncthreads();
  if (!s_last) return;
  for (int w = tid; w < kTailWords; w += 256)
    sk[w] = reinterpret_cast<const unsigned long long*>(t.st)[w];
  double S[4];
  tail_sums_256(t, tid, S);
  if (tid == 0) {
    fin_logic(t.kind, S, reinterpret_cast<KState*>(sk), 0, 0, t.nq, nullptr);
    __hip_atomic_store(t.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  for (int w = tid; w < kTailWords; w += 256)
    reinterpret_cast<unsigned long long*>(t.st)[w] = sk[w];
}

// Folded tail, in place of a finalize launch between the kernel that produced the partial
// sums and the one that needs the scalars: every workgroup of the consumer loads the
// state's scalars into LDS (sk), adds the partial sums in k_finalize's order and runs the
// scalar step on its copy, then reads its scalars from the copy.  The partial sums come
// from an earlier launch, so no fence is needed.  The workgroups hold identical copies and
// the first one writes its copy back (fold_write).  No counter: an arrival count on one
// address serialises a thousand workgroups' atomics.  That is safe when the step writes
// none of the fields it reads (alpha; the half-step test and omega); a step that does
// (the exit test and beta) reads src, a copy of the state that no workgroup of the launch
// writes.
//
// One wavefront of 64 threads: virtual thread v = 64 w + lane of k_finalize's 1024, the
// 16 wave totals added in order.
__device__ __forceinline__ void fold_state_w64(const KTail& f, unsigned long long* sk) {
  const int lane = threadIdx.x & 63;
  const KState* src = f.src ? f.src : f.st;
  for (int w = lane; w < kTailWords; w += 64)
    sk[w] = reinterpret_cast<const unsigned long long*>(src)[w];
  double S[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    S[q] = 0.0;
    if (q >= f.nq || !((f.qmask >> q) & 1)) continue;  // uniform
    const int nb = q < 2 ? f.nblk : f.nblk_hi;
    const double* pq = f.partial + (long)q * kRedBlocks;
    // every load of a round unconditional (a clamped index, the value selected after), so
    // the 16 are in flight together; adding +0.0 for a missing partial keeps the bits
    double vw[16];
#pragma unroll
    for (int w = 0; w < 16; ++w) vw[w] = 0.0;
    const int rounds = (nb + 1023) >> 10;
    for (int r = 0; r < rounds; ++r) {
      double x[16];
#pragma unroll
      for (int w = 0; w < 16; ++w) x[w] = pq[min((r << 10) + (w << 6) + lane, nb - 1)];
#pragma unroll
      for (int w = 0; w < 16; ++w) vw[w] += ((r << 10) + (w << 6) + lane < nb) ? x[w] : 0.0;
    }
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < 16; ++w) s += wave_sum(vw[w]);
    S[q] = s;
  }
  __syncthreads();
  if (lane == 0) fin_logic(f.kind, S, reinterpret_cast<KState*>(sk), 0, 0, f.nq, nullptr);
  __syncthreads();
}

// The same for a workgroup of 256 threads.
__device__ __forceinline__ void fold_state_256(const KTail& f, unsigned long long* sk, int tid) {
  const KState* src = f.src ? f.src : f.st;
  for (int w = tid; w < kTailWords; w += 256)
    sk[w] = reinterpret_cast<const unsigned long long*>(src)[w];
  double S[4];
  tail_sums_256(f, tid, S);
  __syncthreads();
  if (tid == 0) fin_logic(f.kind, S, reinterpret_cast<KState*>(sk), 0, 0, f.nq, nullptr);
  __syncthreads();
}

// The first workgroup of the launch writes its state copy to st (and to copy).
__device__ __forceinline__ void fold_write(const KTail& f, const unsigned long long* sk,
                                           int tid, int nthreads) {
  if (blockIdx.x != 0 || blockIdx.y != 0 || blockIdx.z != 0) return;
  for (int w = tid; w < kTailWords; w += nthreads) {
    reinterpret_cast<unsigned long long*>(f.st)[w] = sk[w];
    if (f.copy) reinterpret_cast<unsigned long long*>(f.copy)[w] = sk[w];
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
// FP32, the dots accumulate in FP64.  x is touched by no other kernel of the iteration, and
// this is the iteration's last read of pHat, rHat and t: those loads and x's store are
// nontemporal (kept out of the caches the other vectors reuse).
template <typename VT>
__device__ __forceinline__ void bicg_upd_elem(long e, VT* __restrict__ x,
                                              const VT* __restrict__ ph,
                                              const VT* __restrict__ rh, VT* __restrict__ r,
                                              const VT* __restrict__ t,
                                              const VT* __restrict__ rt, double alpha,
                                              double omega, double (&acc)[2]) {
  if constexpr (sizeof(VT) == 8) {
    double xv = (double)__builtin_nontemporal_load(x + e);
    xv = xv + alpha * (double)__builtin_nontemporal_load(ph + e);
    xv = xv + omega * (double)__builtin_nontemporal_load(rh + e);
    double rv = (double)r[e] - omega * (double)__builtin_nontemporal_load(t + e);
    __builtin_nontemporal_store((VT)xv, x + e);
    r[e] = (VT)rv;
    acc[0] += rv * rv;
    acc[1] += (double)rt[e] * rv;
  } else {
    float xf = (float)__builtin_nontemporal_load(x + e);
    xf = xf + (float)alpha * (float)__builtin_nontemporal_load(ph + e);
    xf = xf + (float)omega * (float)__builtin_nontemporal_load(rh + e);
    float rf = (float)r[e] - (float)omega * (float)__builtin_nontemporal_load(t + e);
    __builtin_nontemporal_store((VT)xf, x + e);
    r[e] = (VT)rf;
    acc[0] += (double)rf * (double)rf;
    acc[1] += (double)rt[e] * (double)rf;
  }
}

}  // namespace asam
