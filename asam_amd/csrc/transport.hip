// transport.hip -- gfx950 transport stage solve: BiCGStab over the SpDiag transport
// Jacobian with the forward Gauss-Seidel preconditioner (reference
// SRC/MODULE/SpMatrix4CellPar_Mod.f90:56-348, SRC/CG/MatSpDiag_Mod.f90:95-296,
// SRC/MODULE/Vector4Cell_Mod.f90:605-776, 1120-1340).
//
// Data layout in HBM (one block, haloed extents X = nx+2, Y = ny+2, Z = nz+2, m = X*Y*Z):
//   natural: the reference's arrays, x fastest: Val(m, 7) column-major for JacTMom and
//     JacTPot, Val(m, 2) for JacFall; vectors ncomp x m; LU factors ncomp*ncomp x ncell.
//   skewed (Gauss-Seidel only): row iy, step s = ix + iz, level iz:
//     k = (iy*S + s)*Z + iz, S = X + Z - 1.  A wavefront owns one y-row, its 64 lanes the
//     z-levels (iz = lane + 64*slot); at step s every lane updates cell (s - iz, iy, iz),
//     so one step is one contiguous run of Z entries.  The coefficients are stored per
//     row of the matrix (w_d[k] = A(r, r + DiagPtr(d))), so a step reads 7 contiguous
//     runs.  The skew costs (Y*S*Z)/m ~ 1 + Z/X of padding.
//
// The Gauss-Seidel sweep is the reference's lexicographic forward sweep.  Cell (ix,iy,iz)
// needs the new values of (ix-1), (iy-1), (iz-1) and the old ones of (ix+1), (iy+1),
// (iz+1), so every cell of the hyperplane ix+iy+iz = h can be updated at once.  A sweep is
// t = b - U x_old (a parallel pass) followed by the forward substitution, pipelined over
// rows: the wavefront of row iy runs one step behind the one of row iy-1.  Inside a
// workgroup the new values of row iy-1 arrive through LDS (one barrier per step); between
// workgroups through a boundary-row buffer whose entries carry their own readiness (a
// sentinel until written).  Workgroups take their row block from an atomic ticket, so a
// workgroup only ever waits for one that started before it.  Every wait is bounded: past
// the limit the kernel records an error and stops waiting, so every wave reaches the end.
//
// Products and sums are not contracted (no FMA) in the operator, the sweep and the
// vector updates, so they round as the reference's separate multiplies and adds.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "asam_transport.h"
#include "kcommon.h"

namespace asam {
namespace tr {

constexpr int kMaxComp = ASAM_TR_MAX_COMP;
constexpr int kDotBlocks = 1024;
constexpr int kSpinLimit = 1 << 22;

struct Ops {
  int nc;
  int op[kMaxComp];
  double alpha[kMaxComp];
};

struct Dims {
  int nx, ny, nz, X, Y, Z, S;
  long m, ncell, sks;  // sks = Y*S*Z
};

__device__ __forceinline__ bool is_interior(const Dims& d, long r, int& i, int& j, int& k) {
  i = (int)(r % d.X);
  const long t = r / d.X;
  j = (int)(t % d.Y);
  k = (int)(t / d.Y);
  return i >= 1 && i <= d.nx && j >= 1 && j <= d.ny && k >= 1 && k <= d.nz;
}

// ------------------------------------------------------------------ operator (MatVec)
// y = x + LU^-1 (T x - x) on interior cells, T x on the halo, for all components at once
// (MatVecChem/Met_Vec4CellSp: SpAVec, Axpy(-1), PSolveMatrix, Xpy).  The 7 (14 with
// JacTMom) coefficients of a cell are read once for every component.
template <int NC>
__global__ __launch_bounds__(256) void k_matvec(Dims d, const double* __restrict__ tmom,
                                                const double* __restrict__ tpot,
                                                const double* __restrict__ fall,
                                                const double* __restrict__ lu,
                                                const double* __restrict__ x,
                                                double* __restrict__ y, Ops ops, int init,
                                                int use_mom, int use_pot) {
#pragma clang fp contract(off)
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= d.m) return;
  const long XY = (long)d.X * d.Y;
  const long off[7] = {-XY, -(long)d.X, -1, 0, 1, (long)d.X, XY};
  double vp[7], vm[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const long src = r + off[q];
    const bool ok = src >= 0 && src < d.m;
    vp[q] = (use_pot && ok) ? tpot[q * d.m + src] : 0.0;
    vm[q] = (use_mom && ok) ? tmom[q * d.m + src] : 0.0;
  }
  double s[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double* xc = x + c * d.m;
    const int op = init ? ASAM_TR_POT : ops.op[c];
    const double* v = (op == ASAM_TR_MOM) ? vm : vp;
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      const long src = r + off[q];
      if (src >= 0 && src < d.m) acc += v[q] * xc[src];
    }
    if (op == ASAM_TR_FALL) {  // SpABVec: JacFall (offsets 0, XY) scaled by ScaleMat2
      acc += (fall[r] * ops.alpha[c]) * xc[r];
      if (r + XY < d.m) acc += (fall[d.m + r + XY] * ops.alpha[c]) * xc[r + XY];
    }
    s[c] = -1.0 * xc[r] + acc;
  }
  int i, j, k;
  if (lu && is_interior(d, r, i, j, k)) {
    const long q = (i - 1) + (long)d.nx * ((j - 1) + (long)d.ny * (k - 1));
#pragma unroll
    for (int kc = 0; kc < NC; ++kc)
#pragma unroll
      for (int ic = kc + 1; ic < NC; ++ic)
        s[ic] = s[ic] - lu[((long)ic * NC + kc) * d.ncell + q] * s[kc];
#pragma unroll
    for (int kc = NC - 1; kc >= 0; --kc) {
      s[kc] = s[kc] / lu[((long)kc * NC + kc) * d.ncell + q];
#pragma unroll
      for (int ic = 0; ic < kc; ++ic)
        s[ic] = s[ic] - lu[((long)ic * NC + kc) * d.ncell + q] * s[kc];
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) y[c * d.m + r] = x[c * d.m + r] + s[c];
}

// ------------------------------------------------------------------ skewed layout
__device__ __forceinline__ long skew_of(const Dims& d, long r) {
  const int i = (int)(r % d.X);
  const long t = r / d.X;
  const int j = (int)(t % d.Y), k = (int)(t / d.Y);
  return ((long)j * d.S + (i + k)) * d.Z + k;
}

// Row coefficients w_q[k] = A(r, r + DiagPtr(q)) = Val(r + DiagPtr(q), q) of every
// valid skewed slot (padding slots stay 0).
__global__ void k_skew_coef(Dims d, const double* __restrict__ val, double* __restrict__ w) {
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= d.m) return;
  const long XY = (long)d.X * d.Y;
  const long off[7] = {-XY, -(long)d.X, -1, 0, 1, (long)d.X, XY};
  const long k = skew_of(d, r);
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const long src = r + off[q];
    w[q * d.sks + k] = (src >= 0 && src < d.m) ? val[q * d.m + src] : 0.0;
  }
}
// JacFall rows: f0 = its diagonal at r, f1 = its coupling of row r to r + XY.
__global__ void k_skew_fall(Dims d, const double* __restrict__ fall, double* __restrict__ wf) {
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= d.m) return;
  const long XY = (long)d.X * d.Y;
  const long k = skew_of(d, r);
  wf[k] = fall[r];
  wf[d.sks + k] = (r + XY < d.m) ? fall[d.m + r + XY] : 0.0;
}
__global__ void k_skew_vec(Dims d, int nc, const double* __restrict__ v, double* __restrict__ vs) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)nc * d.m) return;
  const long c = e / d.m, r = e - c * d.m;
  vs[c * d.sks + skew_of(d, r)] = v[e];
}
// Back to the natural layout with the six halo faces zero (PSolveSp :734-739).
__global__ void k_unskew(Dims d, int nc, const double* __restrict__ vs, double* __restrict__ v) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)nc * d.m) return;
  const long c = e / d.m, r = e - c * d.m;
  int i, j, k;
  v[e] = is_interior(d, r, i, j, k) ? vs[c * d.sks + skew_of(d, r)] : 0.0;
}

__global__ void k_zero_int(int* p, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = 0;
}

// ------------------------------------------------------------------ Gauss-Seidel sweep
// A sweep is split as the reference's loops are (GaussSeidel_SpDiag :216-234): first
// t = b - U x_old for every cell (k_gs_upart, fully parallel), then the forward
// substitution (L+D) x = t (k_gs_lower, the pipelined part).  Term order of a row as in
// the reference: b, -U (+1, +X, +XY, then JacFall's +XY), -L (-XY, -X, -1), / diag.
constexpr unsigned long long kSentinel = 0xFFF7DEADBEEF0001ull;  // NaN payload: not yet

struct GsArgs {
  Dims d;
  const double* wmom;  // 7 x sks row coefficients (skewed)
  const double* wpot;  // 7 x sks
  const double* wf;    // 2 x sks JacFall rows
  const double* bs;    // nc x sks right-hand side
  const double* ts;    // nc x sks right-hand side of the forward substitution
  double* tw;          // nc x sks: written by k_gs_upart
  double* xs;          // nc x sks: old values in, new values out
  double* bnd;         // nc x G x S x Z: last row of each block, for the next block
  int* counter;        // nc tickets
  int* err;            // set when a wait gives up
  int G;
  int spin_limit;      // polls of a boundary value before a block gives up
  int hold;            // test hook: no block publishes its last row
};

// t = b - U x_old on every valid slot (sweeps after the first), all components of a cell
// in one thread so the U coefficients are read once per cell.
template <int NC>
__global__ __launch_bounds__(256) void k_gs_upart(GsArgs a, Ops ops, int use_mom, int use_pot) {
#pragma clang fp contract(off)
  const Dims& d = a.d;
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.sks) return;
  const int iz = (int)(k % d.Z);
  const long t1 = k / d.Z;
  const int s = (int)(t1 % d.S), iy = (int)(t1 / d.S), ix = s - iz;
  if (ix < 0 || ix >= d.X) return;
  const long SZ = (long)d.S * d.Z;
  const bool ex = ix + 1 < d.X, ey = iy + 1 < d.Y, ez = iz + 1 < d.Z;
  double up[3] = {0.0, 0.0, 0.0}, um[3] = {0.0, 0.0, 0.0};
  if (use_pot) {
    up[0] = a.wpot[4 * d.sks + k], up[1] = a.wpot[5 * d.sks + k], up[2] = a.wpot[6 * d.sks + k];
  }
  if (use_mom) {
    um[0] = a.wmom[4 * d.sks + k], um[1] = a.wmom[5 * d.sks + k], um[2] = a.wmom[6 * d.sks + k];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int op = ops.op[c];
    const double* w = op == ASAM_TR_MOM ? um : up;
    const double* __restrict__ x = a.xs + c * d.sks;
    double t = a.bs[c * d.sks + k];
    if (ex) t = t - w[0] * x[k + d.Z];
    if (ey) t = t - w[1] * x[k + SZ];
    if (ez) {
      t = t - w[2] * x[k + d.Z + 1];
      if (op == ASAM_TR_FALL) t = t - (a.wf[d.sks + k] * ops.alpha[c]) * x[k + d.Z + 1];
    }
    a.tw[c * d.sks + k] = t;
  }
}

__global__ void k_fill_u64(unsigned long long* p, long n, unsigned long long v) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = v;
}

__device__ __forceinline__ bool is_sentinel(double v) {
  return (unsigned long long)__double_as_longlong(v) == kSentinel;
}

// Boundary values of C steps for the 128 z-slots of a row, two per lane (z = lane and
// lane + 64), as one polling wavefront holds them.
template <int C>
struct Pend {
  double v[C][2];
  bool ok;
};

// Re-reads the boundary values that are still the sentinel until every value the first
// row will use is published; bounded, and kept out of line so that the rare wait does not
// bloat the kernel's code.  ok = false when it gave up (the error word is then set).
template <int C>
__device__ __attribute__((noinline)) Pend<C> settle_boundary(Pend<C> p, const double* bnd_in,
                                                           int T0, int S, int Z, int X,
                                                           int limit, int* err) {
  const int lane = threadIdx.x & 63;
  int n = 0;
  p.ok = true;
  while (true) {
    bool miss = false;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int sl = T0 + q;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int z = lane + 64 * h, ix = sl - z;
        const bool need = sl >= 0 && sl < S && z < Z && ix >= 0 && ix < X;
        miss |= need && is_sentinel(p.v[q][h]);
      }
    }
    if (!__any(miss)) return p;
    if (++n > limit) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      p.ok = false;
      return p;
    }
    __builtin_amdgcn_s_sleep(2);
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int sl = T0 + q;
      const long o = (long)(sl < 0 ? 0 : (sl >= S ? S - 1 : sl)) * Z;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int z = lane + 64 * h;
        if (is_sentinel(p.v[q][h]))
          p.v[q][h] = __hip_atomic_load(bnd_in + o + (z < Z ? z : Z - 1), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

// Forward substitution (L+D) x = t, pipelined over rows.  A workgroup is R rows of
// Q = ceil(Z/64) compute wavefronts (lane iz = 64*qz + lane) plus one polling wavefront;
// step T of the workgroup is step s = T - j of row j (row iy = g*R + j).
//  - The steps run in chunks of C.  Each compute wavefront keeps the t and L/D
//    coefficients of K chunks in registers and refills a chunk's registers as soon as its
//    steps are done, so the loads of a step are issued (K-1)*C steps before it runs.
//  - New values of row iy-1 come through LDS (one barrier per step); the z-neighbour from
//    the lane below (through LDS across the two wavefronts of a row), the x-neighbour
//    from the lane's own previous step.
//  - The first row's neighbour row is the previous block's last row.  Its values are
//    written with agent-scope atomic stores into a buffer pre-filled with a NaN-payload
//    sentinel; a value that is not the sentinel is final, so no fence or flag orders
//    them.  The polling wavefront reads them with agent-scope atomic loads, a chunk
//    ahead and in a wavefront of their own (so no coefficient load is queued ahead of a
//    poll), and hands them to the first row through LDS.
//  - Blocks take their rows from an atomic ticket and only wait for a block that started
//    before them; every wait is bounded.
template <int C>
struct Chunk {
  double t[C], w0[C], w1[C], w2[C], w3[C], f0[C];
};

constexpr int kBoundStride = 128;  // z-stride of the boundary-value slots in LDS

template <int C, int K, bool FALL>
__global__ __launch_bounds__(576) void k_gs_lower(GsArgs a, Ops ops, int Q) {
#pragma clang fp contract(off)
  extern __shared__ double lds[];  // [2][R][Z] row exchange, then [2][C][128] boundary values
  __shared__ int s_ticket;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nw = (blockDim.x >> 6) - 1;  // compute wavefronts; the last one polls
  const bool poller = wv == nw;
  const int R = nw / Q;
  const int j = poller ? 0 : wv / Q, qz = poller ? 0 : wv - j * Q;
  const Dims& d = a.d;
  const int c = blockIdx.x / a.G;
  const int op = ops.op[c];
  // a launch serves the components of one kind (with JacFall or not); the others leave
  // at once, whole workgroups together, before any barrier
  if ((op == ASAM_TR_FALL) != FALL) return;
  if (threadIdx.x == 0) s_ticket = atomicAdd(a.counter + c, 1);
  __syncthreads();
  const int g = s_ticket;
  const int S = d.S, Z = d.Z, X = d.X;
  const long SZ = (long)S * Z;
  const int steps = S + R - 1;
  // every wavefront runs whole loop turns of K chunks: the steps past `steps` compute
  // nothing (no cell is valid there), and no branch skips part of a turn, so the
  // compiler's load-wait counts stay exact across the turn instead of draining
  const int nchunks = (steps + K * C - 1) / (K * C) * K;
  double* byL = lds + 2 * R * Z;
  auto slot = [&](int sl) { return sl < 0 ? 0 : (sl >= S ? S - 1 : sl); };

  if (poller) {
    // every barrier of the compute wavefronts is matched here, one per step
    const bool active = g > 0;
    const double* bnd_in = a.bnd + ((long)c * a.G + (g > 0 ? g - 1 : 0)) * SZ;
    Pend<C> p;
    bool broken = false;
    auto issue = [&](int T0) {
#pragma unroll
      for (int q = 0; q < C; ++q) {
        const long o = (long)slot(T0 + q) * Z;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int z = lane + 64 * h;
          p.v[q][h] = __hip_atomic_load(bnd_in + o + (z < Z ? z : Z - 1), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    };
    // waits for the values of the chunk at T0 and puts them in LDS slot kc & 1
    auto settle = [&](int T0, int kc) {
      bool miss = false;
#pragma unroll
      for (int q = 0; q < C; ++q) {
        const int sl = T0 + q;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int z = lane + 64 * h, ix = sl - z;
          miss |= sl >= 0 && sl < S && z < Z && ix >= 0 && ix < X && is_sentinel(p.v[q][h]);
        }
      }
      if (__any(miss) && !broken) {
        p = settle_boundary<C>(p, bnd_in, T0, S, Z, X, a.spin_limit, a.err);
        broken = !p.ok;
      }
#pragma unroll
      for (int q = 0; q < C; ++q)
#pragma unroll
        for (int h = 0; h < 2; ++h)
          byL[((kc & 1) * C + q) * kBoundStride + lane + 64 * h] = p.v[q][h];
    };
    if (active) {
      issue(0);
      settle(0, 0);
    }
    __syncthreads();
    for (int kc = 0; kc < nchunks; ++kc) {
      const int T0 = kc * C;
#pragma unroll
      for (int q = 0; q < C; ++q) {
        // the next chunk's values: requested after this chunk's first barrier, in LDS
        // before its last one
        if (active && q == 0 && T0 + C < steps) issue(T0 + C);
        if (active && q == C - 1 && T0 + C < steps) settle(T0 + C, kc + 1);
        __syncthreads();
      }
    }
    return;
  }

  const int iy = g * R + j;
  const bool row_ok = iy < d.Y;
  const double alpha = ops.alpha[c];
  const double* __restrict__ W = (op == ASAM_TR_MOM) ? a.wmom : a.wpot;
  const long sks = d.sks;
  const double* __restrict__ ts = a.ts + c * sks;
  double* xs = a.xs + c * sks;
  const long rowbase = (long)(row_ok ? iy : 0) * SZ;
  const int iz = 64 * qz + lane;
  const int izc = iz < Z ? iz : Z - 1;
  const bool publish = j == R - 1 && g + 1 < a.G && row_ok && !a.hold;
  double* bnd_out = a.bnd + ((long)c * a.G + g) * SZ;

  auto load = [&](int T0, Chunk<C>& ch) {
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const long k = rowbase + (long)slot(T0 + q - j) * Z + izc;
      ch.t[q] = ts[k];
      ch.w0[q] = W[k];
      ch.w1[q] = W[sks + k];
      ch.w2[q] = W[2 * sks + k];
      ch.w3[q] = W[3 * sks + k];
      ch.f0[q] = FALL ? a.wf[k] : 0.0;
    }
  };
  double prev = 0.0;
  // runs the C steps from T0 with coefficients ch, then refills ch with the coefficients
  // of the chunk K chunks later
  auto run = [&](int T0, Chunk<C>& ch) {
    const int kc = T0 / C;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int T = T0 + q;
      const int s = T - j, ix = s - iz;
      const bool valid = row_ok && s >= 0 && s < S && iz < Z && ix >= 0 && ix < X;
      const double below = __shfl(prev, (lane + 63) & 63);
      const int pb = ((T - 1) & 1) * R * Z, cb = (T & 1) * R * Z;
      // row iy-1: the previous step of row j-1, or for the first row the boundary value
      double xy_l = 0.0, xz_l = 0.0;
      if (iz < Z)
        xy_l = j > 0 ? lds[pb + (j - 1) * Z + iz] : byL[((kc & 1) * C + q) * kBoundStride + iz];
      if (lane == 0 && qz > 0) xz_l = lds[pb + j * Z + iz - 1];
      double out = 0.0;
      if (valid) {
        double t = ch.t[q];
        if (iz > 0) t = t - ch.w0[q] * (lane > 0 ? below : xz_l);
        if (iy > 0) t = t - ch.w1[q] * xy_l;
        if (ix > 0) t = t - ch.w2[q] * prev;
        double dg = ch.w3[q];
        if (FALL) dg = dg + ch.f0[q] * alpha;
        out = t / dg;
        xs[rowbase + (long)s * Z + iz] = out;
        if (publish)
          __hip_atomic_store(bnd_out + (long)s * Z + iz, out, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      }
      if (iz < Z) lds[cb + j * Z + iz] = out;
      prev = out;
      __syncthreads();
    }
    load(T0 + K * C, ch);
  };
  Chunk<C> buf[K];
#pragma unroll
  for (int b = 0; b < K; ++b) load(b * C, buf[b]);
  __syncthreads();  // the first chunk's boundary values are in LDS
  for (int T0 = 0; T0 < nchunks * C; T0 += K * C) {
#pragma unroll
    for (int b = 0; b < K; ++b) run(T0 + b * C, buf[b]);
  }
}

// ------------------------------------------------------------------ dots and updates
// Sums over the interior cells of all components of NQ products a_q . b_q.
template <int NQ>
__global__ __launch_bounds__(256) void k_dot(Dims d, int nc, const double* __restrict__ a0,
                                             const double* __restrict__ b0,
                                             const double* __restrict__ a1,
                                             const double* __restrict__ b1,
                                             double* __restrict__ partial) {
  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
  const long total = (long)nc * d.ncell;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const long c = e / d.ncell, q = e - c * d.ncell;
    const long i = q % d.nx, t = q / d.nx;
    const long jj = t % d.ny, k = t / d.ny;
    const long r = c * d.m + (i + 1) + d.X * ((jj + 1) + (long)d.Y * (k + 1));
    acc[0] += a0[r] * b0[r];
    if (NQ > 1) acc[NQ - 1] += a1[r] * b1[r];
  }
  block_reduce_store<NQ>(acc, partial, blockIdx.x, threadIdx.x, blockDim.x, kDotBlocks);
}
// Fixed-order sum of the partials: thread t adds partials t, t+1024, ... (one each here),
// then the block tree; every run gives the same bits.
template <int NQ>
__global__ __launch_bounds__(1024) void k_dot_fin(const double* __restrict__ partial, int nblk,
                                                  double* __restrict__ out) {
  double v[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += partial[q * kDotBlocks + b];
    v[q] = s;
  }
  block_reduce_store<NQ>(v, out, 0, threadIdx.x, blockDim.x, 1);
}

__global__ void k_axpy(long n, double a, const double* __restrict__ x, double* __restrict__ y) {
#pragma clang fp contract(off)
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) y[e] = a * x[e] + y[e];
}
__global__ void k_xpay(long n, const double* __restrict__ x, double a, double* __restrict__ y) {
#pragma clang fp contract(off)
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) y[e] = a * y[e] + x[e];
}

}  // namespace tr
}  // namespace asam

using namespace asam;
using namespace asam::tr;

struct asam_tr_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  Dims d{};
  int nc = 0, iterg = 5, M = 1;
  Ops ops_mv{}, ops_ps{};  // per-component operator of MatVec and of PSolve
  asam_tr_tuning tune{8, 2, 0, 0};
  int hold_boundary = 0;  // test hook (asam_tr_test_hold_boundary)
  bool have_mom = false, have_pot = false, have_fall = false, have_lu = false;
  std::string err;
  std::vector<void*> allocs;
  double *tmom = nullptr, *tpot = nullptr, *fall = nullptr, *lu = nullptr;
  double *wmom = nullptr, *wpot = nullptr, *wf = nullptr;
  double *bs = nullptr, *xs = nullptr, *ts = nullptr, *bnd = nullptr;
  double *vx = nullptr, *vb = nullptr, *vr = nullptr, *vrt = nullptr, *vrh = nullptr;
  double *vp = nullptr, *vph = nullptr, *vt = nullptr, *vv = nullptr, *vh = nullptr;
  double *partial = nullptr, *dots = nullptr;
  int *sync = nullptr;  // tickets, error word
  int sync_n = 0;
  double* host_dots = nullptr;  // pinned
};

#define TRCHK(ctx, expr)                                                   \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess) {                                                \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);      \
      return ASAM_E_HIP;                                                   \
    }                                                                      \
  } while (0)

static int trfail(asam_tr_ctx* c, int code, const char* msg) {
  if (c) c->err = msg;
  return code;
}

template <typename T>
static int tralloc(asam_tr_ctx* c, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, (count ? count : 1) * sizeof(T));
  if (e != hipSuccess) {
    c->err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return ASAM_E_NOMEM;
  }
  c->allocs.push_back(q);
  *p = (T*)q;
  return 0;
}

static inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Rows per workgroup: at most 8 compute wavefronts (M per row, M = 2 when Z > 64), plus
// the one that polls the previous block's boundary row.
static int rows_per_wg(const asam_tr_ctx* c) {
  int R = c->tune.gs_rows;
  if (R < 1) R = 1;
  if (R > 8 / c->M) R = 8 / c->M;
  return R;
}
static int blocks_y(const asam_tr_ctx* c) {
  const int R = rows_per_wg(c);
  return (c->d.Y + R - 1) / R;
}

extern "C" const char* asam_tr_last_error(const asam_tr_ctx* c) {
  return c ? c->err.c_str() : "null context";
}

extern "C" int asam_tr_destroy(asam_tr_ctx* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->host_dots) (void)hipHostFree(c->host_dots);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

extern "C" int asam_tr_create(asam_tr_ctx** out, const asam_tr_grid* g) {
  if (!out || !g) return ASAM_E_INVALID;
  *out = nullptr;
  if (g->nx < 1 || g->ny < 1 || g->nz < 1 || g->nz > 126 || g->ncomp < 1 ||
      g->ncomp > kMaxComp || g->iterg < 0 || g->iterg > 32)
    return ASAM_E_INVALID;
  const long X = g->nx + 2L, Y = g->ny + 2L, Z = g->nz + 2L;
  if (X * Y * Z * g->ncomp > (1L << 31)) return ASAM_E_INVALID;
  asam_tr_ctx* c = new asam_tr_ctx();
  auto bail = [&](int rc) {
    asam_tr_destroy(c);
    return rc;
  };
  if (hipGetDevice(&c->device) != hipSuccess) return bail(ASAM_E_HIP);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(ASAM_E_HIP);
  Dims& d = c->d;
  d.nx = g->nx, d.ny = g->ny, d.nz = g->nz;
  d.X = (int)X, d.Y = (int)Y, d.Z = (int)Z, d.S = (int)(X + Z - 1);
  d.m = X * Y * Z;
  d.ncell = (long)g->nx * g->ny * g->nz;
  d.sks = Y * d.S * Z;
  c->nc = g->ncomp;
  c->iterg = g->iterg;
  c->M = Z <= 64 ? 1 : 2;
  c->ops_mv.nc = c->ops_ps.nc = c->nc;
  for (int q = 0; q < kMaxComp; ++q) {
    c->ops_mv.op[q] = c->ops_ps.op[q] = ASAM_TR_POT;
    c->ops_mv.alpha[q] = c->ops_ps.alpha[q] = 0.0;
  }
  const size_t V = (size_t)c->nc * d.m;
  int rc = 0;
  rc |= tralloc(c, &c->tmom, 7 * (size_t)d.m);
  rc |= tralloc(c, &c->tpot, 7 * (size_t)d.m);
  rc |= tralloc(c, &c->fall, 2 * (size_t)d.m);
  rc |= tralloc(c, &c->lu, (size_t)c->nc * c->nc * d.ncell);
  rc |= tralloc(c, &c->wmom, 7 * (size_t)d.sks);
  rc |= tralloc(c, &c->wpot, 7 * (size_t)d.sks);
  rc |= tralloc(c, &c->wf, 2 * (size_t)d.sks);
  rc |= tralloc(c, &c->bs, (size_t)c->nc * d.sks);
  rc |= tralloc(c, &c->xs, (size_t)c->nc * d.sks + (size_t)d.S * d.Z + d.Z + 1);
  rc |= tralloc(c, &c->ts, (size_t)c->nc * d.sks);
  rc |= tralloc(c, &c->bnd, (size_t)c->nc * d.sks);
  double** vecs[] = {&c->vx, &c->vb, &c->vr, &c->vrt, &c->vrh, &c->vp, &c->vph, &c->vt, &c->vv,
                     &c->vh};
  for (double** v : vecs) rc |= tralloc(c, v, V);
  rc |= tralloc(c, &c->partial, 2 * (size_t)kDotBlocks);
  rc |= tralloc(c, &c->dots, 2);
  c->sync_n = c->nc + 1;  // tickets, error word
  rc |= tralloc(c, &c->sync, (size_t)c->sync_n);
  if (rc) return bail(ASAM_E_NOMEM);
  if (hipHostMalloc((void**)&c->host_dots, 2 * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return bail(ASAM_E_NOMEM);
  // skewed padding slots must read as 0
  if (hipMemsetAsync(c->wmom, 0, 7 * sizeof(double) * d.sks, c->stream) != hipSuccess ||
      hipMemsetAsync(c->wpot, 0, 7 * sizeof(double) * d.sks, c->stream) != hipSuccess ||
      hipMemsetAsync(c->wf, 0, 2 * sizeof(double) * d.sks, c->stream) != hipSuccess ||
      hipMemsetAsync(c->xs, 0, sizeof(double) * ((size_t)c->nc * d.sks + (size_t)d.S * d.Z + d.Z + 1),
                     c->stream) != hipSuccess ||
      hipMemsetAsync(c->bs, 0, (size_t)c->nc * sizeof(double) * d.sks, c->stream) != hipSuccess ||
      hipMemsetAsync(c->sync, 0, sizeof(int) * c->sync_n, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return bail(ASAM_E_HIP);
  *out = c;
  return 0;
}

extern "C" int asam_tr_set_matrix(asam_tr_ctx* c, int which, int numdiag, int diag,
                                  const int* ptr, const double* val) {
  if (!c || !ptr || !val) return trfail(c, ASAM_E_INVALID, "null argument");
  const Dims& d = c->d;
  const long X = d.X, XY = (long)d.X * d.Y, m = d.m;
  if (which == ASAM_TR_JACTMOM || which == ASAM_TR_JACTPOT) {
    const long want[7] = {-XY, -X, -1, 0, 1, X, XY};
    if (numdiag != 7 || diag != 4) return trfail(c, ASAM_E_INVALID, "NumDiag must be 7, Diag 4");
    for (int q = 0; q < 7; ++q)
      if (ptr[q] != want[q]) return trfail(c, ASAM_E_INVALID, "DiagPtr is not the 7-point layout");
    // couplings across the wrap of the linear index are not neighbour couplings
    for (long r = 0; r < m; ++r) {
      const long i = r % X, jy = (r / X) % d.Y;
      if ((i == X - 1 && val[2 * m + r] != 0.0) || (i == 0 && r > 0 && val[4 * m + r] != 0.0) ||
          (jy == d.Y - 1 && val[1 * m + r] != 0.0) || (jy == 0 && r >= X && val[5 * m + r] != 0.0))
        return trfail(c, ASAM_E_INVALID,
                      "coupling between the ends of consecutive x-rows or y-planes");
      for (int q = 0; q < 7; ++q)
        if (!std::isfinite(val[q * m + r])) return trfail(c, ASAM_E_NAN, "non-finite Val");
    }
    double* dst = which == ASAM_TR_JACTMOM ? c->tmom : c->tpot;
    double* w = which == ASAM_TR_JACTMOM ? c->wmom : c->wpot;
    TRCHK(c, hipSetDevice(c->device));
    TRCHK(c, hipMemcpyAsync(dst, val, 7 * sizeof(double) * m, hipMemcpyHostToDevice, c->stream));
    k_skew_coef<<<blocks_for(m, 256), 256, 0, c->stream>>>(d, dst, w);
    TRCHK(c, hipGetLastError());
    TRCHK(c, hipStreamSynchronize(c->stream));
    (which == ASAM_TR_JACTMOM ? c->have_mom : c->have_pot) = true;
    return 0;
  }
  if (which == ASAM_TR_JACFALL) {
    if (numdiag != 2 || diag != 1 || ptr[0] != 0 || ptr[1] != XY)
      return trfail(c, ASAM_E_INVALID, "JacFall must have DiagPtr 0, (nx+2)(ny+2), Diag 1");
    for (long e = 0; e < 2 * m; ++e)
      if (!std::isfinite(val[e])) return trfail(c, ASAM_E_NAN, "non-finite Val");
    TRCHK(c, hipSetDevice(c->device));
    TRCHK(c, hipMemcpyAsync(c->fall, val, 2 * sizeof(double) * m, hipMemcpyHostToDevice, c->stream));
    k_skew_fall<<<blocks_for(m, 256), 256, 0, c->stream>>>(d, c->fall, c->wf);
    TRCHK(c, hipGetLastError());
    TRCHK(c, hipStreamSynchronize(c->stream));
    c->have_fall = true;
    return 0;
  }
  return trfail(c, ASAM_E_INVALID, "which");
}

extern "C" int asam_tr_set_lu(asam_tr_ctx* c, const double* lu) {
  if (!c) return ASAM_E_INVALID;
  if (!lu) {
    c->have_lu = false;
    return 0;
  }
  const size_t n = (size_t)c->nc * c->nc * c->d.ncell;
  for (size_t e = 0; e < n; ++e)
    if (!std::isfinite(lu[e])) return trfail(c, ASAM_E_NAN, "non-finite LU");
  TRCHK(c, hipSetDevice(c->device));
  TRCHK(c, hipMemcpyAsync(c->lu, lu, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  TRCHK(c, hipStreamSynchronize(c->stream));
  c->have_lu = true;
  return 0;
}

extern "C" int asam_tr_set_components(asam_tr_ctx* c, const int* op_mv, const int* op_ps,
                                      const double* alpha) {
  if (!c) return ASAM_E_INVALID;
  for (int q = 0; q < c->nc; ++q) {
    const int a = op_mv ? op_mv[q] : ASAM_TR_POT, b = op_ps ? op_ps[q] : ASAM_TR_POT;
    if (a < 0 || a > 2 || b < 0 || b > 2) return trfail(c, ASAM_E_INVALID, "operator code");
    if (alpha && !std::isfinite(alpha[q])) return trfail(c, ASAM_E_NAN, "alpha");
  }
  for (int q = 0; q < c->nc; ++q) {
    c->ops_mv.op[q] = op_mv ? op_mv[q] : ASAM_TR_POT;
    c->ops_ps.op[q] = op_ps ? op_ps[q] : ASAM_TR_POT;
    c->ops_mv.alpha[q] = c->ops_ps.alpha[q] = alpha ? alpha[q] : 0.0;
  }
  return 0;
}

// ------------------------------------------------------------------ launchers
static bool uses(const Ops& o, int nc, int op) {
  for (int q = 0; q < nc; ++q)
    if (o.op[q] == op) return true;
  return false;
}

// The matrices a component table needs have been set (JacTPot always: the first
// residual of BiCGStab takes it for every component).
static int check_ready(asam_tr_ctx* c, const Ops& o) {
  if (!c->have_pot) return trfail(c, ASAM_E_STATE, "JacTPot not set");
  if (uses(o, c->nc, ASAM_TR_MOM) && !c->have_mom)
    return trfail(c, ASAM_E_STATE, "JacTMom not set");
  if (uses(o, c->nc, ASAM_TR_FALL) && !c->have_fall)
    return trfail(c, ASAM_E_STATE, "JacFall not set");
  return 0;
}

template <int NC>
static void matvec_nc(asam_tr_ctx* c, const double* x, double* y, int init, int mom) {
  const Dims& d = c->d;
  k_matvec<NC><<<blocks_for(d.m, 256), 256, 0, c->stream>>>(
      d, c->tmom, c->tpot, c->fall, c->have_lu ? c->lu : nullptr, x, y, c->ops_mv, init, mom,
      1);
}

static int launch_matvec(asam_tr_ctx* c, const double* x, double* y, int init) {
  const int mom = (!init && uses(c->ops_mv, c->nc, ASAM_TR_MOM)) ? 1 : 0;
  switch (c->nc) {
#define NCCASE(n) \
  case n:         \
    matvec_nc<n>(c, x, y, init, mom); \
    break;
    NCCASE(1) NCCASE(2) NCCASE(3) NCCASE(4) NCCASE(5) NCCASE(6) NCCASE(7) NCCASE(8)
    NCCASE(9) NCCASE(10) NCCASE(11) NCCASE(12) NCCASE(13) NCCASE(14) NCCASE(15) NCCASE(16)
#undef NCCASE
    default:
      return trfail(c, ASAM_E_INVALID, "ncomp");
  }
  TRCHK(c, hipGetLastError());
  return 0;
}

template <int C, int K, bool FALL>
static void lower_launch(asam_tr_ctx* c, const GsArgs& a, int R, int Q) {
  const size_t lds = (2 * (size_t)R * c->d.Z + 2 * (size_t)C * kBoundStride) * sizeof(double);
  k_gs_lower<C, K, FALL><<<(unsigned)(c->nc * a.G), 64 * (R * Q + 1), lds, c->stream>>>(
      a, c->ops_ps, Q);
}

// gs_chunk 2: chunks of 2 steps, 4 in registers; 4: chunks of 4 steps, 2 in registers
template <bool FALL>
static void lower_kind(asam_tr_ctx* c, const GsArgs& a, int R) {
  if (c->tune.gs_chunk >= 4) lower_launch<4, 2, FALL>(c, a, R, c->M);
  else lower_launch<2, 4, FALL>(c, a, R, c->M);
}

template <int NC>
static void upart_nc(asam_tr_ctx* c, const GsArgs& a, int use_mom, int use_pot) {
  k_gs_upart<NC><<<blocks_for(c->d.sks, 256), 256, 0, c->stream>>>(a, c->ops_ps, use_mom,
                                                                     use_pot);
}

static int launch_upart(asam_tr_ctx* c, const GsArgs& a) {
  int mom = 0, pot = 0;
  for (int q = 0; q < c->nc; ++q) (c->ops_ps.op[q] == ASAM_TR_MOM ? mom : pot) = 1;
  switch (c->nc) {
#define NCCASE(n) \
  case n:         \
    upart_nc<n>(c, a, mom, pot); \
    break;
    NCCASE(1) NCCASE(2) NCCASE(3) NCCASE(4) NCCASE(5) NCCASE(6) NCCASE(7) NCCASE(8)
    NCCASE(9) NCCASE(10) NCCASE(11) NCCASE(12) NCCASE(13) NCCASE(14) NCCASE(15) NCCASE(16)
#undef NCCASE
    default:
      return trfail(c, ASAM_E_INVALID, "ncomp");
  }
  return 0;
}

// One Gauss-Seidel sweep over every component: t = b - U x_old (not for the first
// sweep, which starts from x = 0), then the pipelined forward substitution.  Tickets are
// cleared and the boundary rows reset to the sentinel first.
static int launch_sweep(asam_tr_ctx* c, int sweep) {
  const Dims& d = c->d;
  const int R = rows_per_wg(c), G = blocks_y(c);
  bool any_fall = false, any_plain = false;
  for (int q = 0; q < c->nc; ++q)
    (c->ops_ps.op[q] == ASAM_TR_FALL ? any_fall : any_plain) = true;
  GsArgs a;
  a.d = d;
  a.wmom = c->wmom, a.wpot = c->wpot, a.wf = c->wf;
  a.bs = c->bs, a.tw = c->ts, a.xs = c->xs, a.bnd = c->bnd;
  a.ts = sweep > 0 ? c->ts : c->bs;
  a.counter = c->sync;
  a.err = c->sync + c->sync_n - 1;
  a.G = G;
  a.spin_limit = c->tune.gs_spin_limit > 0 ? c->tune.gs_spin_limit : kSpinLimit;
  a.hold = c->hold_boundary;
  if (sweep > 0)
    if (int rc = launch_upart(c, a)) return rc;
  k_zero_int<<<blocks_for(c->nc, 256), 256, 0, c->stream>>>(c->sync, c->nc);
  const long nb = (long)c->nc * G * d.S * d.Z;
  k_fill_u64<<<blocks_for(nb, 256), 256, 0, c->stream>>>((unsigned long long*)c->bnd, nb,
                                                          kSentinel);
  if (any_plain) lower_kind<false>(c, a, R);
  if (any_fall) lower_kind<true>(c, a, R);
  TRCHK(c, hipGetLastError());
  return 0;
}

// PSolve_Vec4CellSp: y = GS^(1+IterG)(x) per component, halo zero.  x may equal y.
static int launch_psolve(asam_tr_ctx* c, const double* x, double* y) {
  const Dims& d = c->d;
  const long V = (long)c->nc * d.m;
  k_zero_int<<<1, 64, 0, c->stream>>>(c->sync + c->sync_n - 1, 1);
  k_skew_vec<<<blocks_for(V, 256), 256, 0, c->stream>>>(d, c->nc, x, c->bs);
  TRCHK(c, hipGetLastError());
  for (int sw = 0; sw <= c->iterg; ++sw)
    if (int rc = launch_sweep(c, sw)) return rc;
  k_unskew<<<blocks_for(V, 256), 256, 0, c->stream>>>(d, c->nc, c->xs, y);
  TRCHK(c, hipGetLastError());
  return 0;
}

static int check_gs_error(asam_tr_ctx* c) {
  int e = 0;
  TRCHK(c, hipMemcpyAsync(&e, c->sync + c->sync_n - 1, sizeof(int), hipMemcpyDeviceToHost,
                          c->stream));
  TRCHK(c, hipStreamSynchronize(c->stream));
  if (e) return trfail(c, ASAM_E_HIP, "Gauss-Seidel sweep: a workgroup waited past its limit");
  return 0;
}

// NQ dots over the interior cells; results on the host.
template <int NQ>
static int dots(asam_tr_ctx* c, const double* a0, const double* b0, const double* a1,
                const double* b1, double* out) {
  k_dot<NQ><<<kDotBlocks, 256, 0, c->stream>>>(c->d, c->nc, a0, b0, a1, b1, c->partial);
  k_dot_fin<NQ><<<1, 1024, 0, c->stream>>>(c->partial, kDotBlocks, c->dots);
  TRCHK(c, hipGetLastError());
  TRCHK(c, hipMemcpyAsync(c->host_dots, c->dots, NQ * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  TRCHK(c, hipStreamSynchronize(c->stream));
  for (int q = 0; q < NQ; ++q) out[q] = c->host_dots[q];
  return 0;
}

static int axpy(asam_tr_ctx* c, double a, const double* x, double* y) {
  if (a == 0.0) return 0;  // Axpy_VectorCell skips alpha == 0 (:1243)
  const long V = (long)c->nc * c->d.m;
  k_axpy<<<blocks_for(V, 256), 256, 0, c->stream>>>(V, a, x, y);
  TRCHK(c, hipGetLastError());
  return 0;
}
static int xpay(asam_tr_ctx* c, const double* x, double a, double* y) {
  const long V = (long)c->nc * c->d.m;
  k_xpay<<<blocks_for(V, 256), 256, 0, c->stream>>>(V, x, a, y);
  TRCHK(c, hipGetLastError());
  return 0;
}
static int copy(asam_tr_ctx* c, const double* x, double* y) {
  TRCHK(c, hipMemcpyAsync(y, x, sizeof(double) * c->nc * c->d.m, hipMemcpyDeviceToDevice,
                          c->stream));
  return 0;
}

#define TRY(expr)              \
  do {                         \
    if (int rc_ = (expr)) return rc_; \
  } while (0)

// BiCGStabChem/Met_Vec4CellSp (SpMatrix4CellPar_Mod.f90:140-348) on device vectors.
static int solve_dev(asam_tr_ctx* c, double* x, const double* b, int* maxiter, double* tol,
                     asam_tr_info* info) {
  asam_tr_info inf{};
  double *r = c->vr, *rt = c->vrt, *rh = c->vrh, *p = c->vp, *ph = c->vph, *t = c->vt,
         *v = c->vv;
  const int MaxIter = *maxiter;
  const double Tol = *tol;
  TRY(launch_matvec(c, x, r, 1));  // MatVecChem(r, A, x)
  inf.matvecs = 1;
  TRY(xpay(c, b, -1.0, r));        // r = b - r
  TRY(copy(c, r, rt));
  double q2[2];
  TRY(dots<1>(c, r, r, nullptr, nullptr, q2));
  const double rnorm0 = std::sqrt(q2[0]);
  inf.rnorm0 = rnorm0;
  inf.rnorm = rnorm0;
  inf.status = ASAM_NOT_CONVERGED;
  double rho = 0.0, rho_old, alpha = 0.0, omega = 0.0, beta, rnorm = rnorm0;
  for (int it = 1; it <= MaxIter; ++it) {
    inf.iterations = it;
    rho_old = rho;
    TRY(dots<1>(c, rt, r, nullptr, nullptr, q2));
    rho = q2[0];
    if (rho == 0.0) {
      *maxiter = it;
      *tol = 0.0;
      inf.status = ASAM_BREAKDOWN;
      break;
    }
    if (it == 1) {
      TRY(copy(c, r, p));
    } else {
      beta = (rho / rho_old) * (alpha / omega);
      TRY(axpy(c, -omega, v, p));
      TRY(xpay(c, r, beta, p));
    }
    TRY(launch_psolve(c, p, ph));
    TRY(launch_matvec(c, ph, v, 0));
    inf.psolves++, inf.matvecs++;
    TRY(dots<1>(c, rt, v, nullptr, nullptr, q2));
    alpha = rho / q2[0];
    TRY(axpy(c, -alpha, v, r));
    TRY(dots<1>(c, r, r, nullptr, nullptr, q2));
    rnorm = std::sqrt(q2[0]);
    if (rnorm <= Tol * rnorm0) {
      TRY(axpy(c, alpha, ph, x));
      *maxiter = -it;
      *tol = rnorm;
      inf.status = ASAM_CONVERGED;
      break;
    }
    TRY(launch_psolve(c, r, rh));
    TRY(launch_matvec(c, rh, t, 0));
    inf.psolves++, inf.matvecs++;
    TRY(dots<2>(c, t, r, t, t, q2));
    omega = q2[0] / q2[1];
    TRY(axpy(c, alpha, ph, x));
    TRY(axpy(c, omega, rh, x));
    TRY(axpy(c, -omega, t, r));
    TRY(dots<1>(c, r, r, nullptr, nullptr, q2));
    rnorm = std::sqrt(q2[0]);
    if (rnorm <= Tol * rnorm0) {
      *maxiter = it;
      *tol = rnorm;
      inf.status = ASAM_CONVERGED;
      break;
    }
    if (it == MaxIter) {
      *maxiter = 2 * it;
      *tol = rnorm;
      break;
    }
  }
  inf.rnorm = (inf.status == ASAM_BREAKDOWN) ? rnorm : *tol;
  if (!std::isfinite(rnorm)) return trfail(c, ASAM_E_NAN, "non-finite residual");
  TRY(check_gs_error(c));
  if (info) *info = inf;
  return 0;
}

extern "C" int asam_tr_matvec_device(asam_tr_ctx* c, const double* x, double* y, int init) {
  if (!c || !x || !y || x == y) return trfail(c, ASAM_E_INVALID, "x, y must be distinct buffers");
  TRY(check_ready(c, c->ops_mv));
  TRCHK(c, hipSetDevice(c->device));
  TRY(launch_matvec(c, x, y, init));
  TRCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int asam_tr_psolve_device(asam_tr_ctx* c, const double* x, double* y) {
  if (!c || !x || !y) return trfail(c, ASAM_E_INVALID, "null argument");
  TRY(check_ready(c, c->ops_ps));
  TRCHK(c, hipSetDevice(c->device));
  TRY(launch_psolve(c, x, y));
  return check_gs_error(c);
}

extern "C" int asam_tr_solve_device(asam_tr_ctx* c, double* x, const double* b, int* maxiter,
                                    double* tol, asam_tr_info* info) {
  if (!c || !x || !b || !maxiter || !tol) return trfail(c, ASAM_E_INVALID, "null argument");
  if (x == b) return trfail(c, ASAM_E_INVALID, "x and b must be distinct buffers");
  if (!(*tol >= 0.0)) return trfail(c, ASAM_E_INVALID, "tol");
  TRY(check_ready(c, c->ops_ps));
  TRY(check_ready(c, c->ops_mv));
  TRCHK(c, hipSetDevice(c->device));
  return solve_dev(c, x, b, maxiter, tol, info);
}

static size_t vbytes(const asam_tr_ctx* c) { return sizeof(double) * c->nc * c->d.m; }

extern "C" int asam_tr_matvec(asam_tr_ctx* c, const double* x, double* y, int init) {
  if (!c || !x || !y) return trfail(c, ASAM_E_INVALID, "null argument");
  TRY(check_ready(c, c->ops_mv));
  TRCHK(c, hipSetDevice(c->device));
  TRCHK(c, hipMemcpyAsync(c->vh, x, vbytes(c), hipMemcpyHostToDevice, c->stream));
  TRY(launch_matvec(c, c->vh, c->vt, init));
  TRCHK(c, hipMemcpyAsync(y, c->vt, vbytes(c), hipMemcpyDeviceToHost, c->stream));
  TRCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int asam_tr_psolve(asam_tr_ctx* c, const double* x, double* y) {
  if (!c || !x || !y) return trfail(c, ASAM_E_INVALID, "null argument");
  TRY(check_ready(c, c->ops_ps));
  TRCHK(c, hipSetDevice(c->device));
  TRCHK(c, hipMemcpyAsync(c->vh, x, vbytes(c), hipMemcpyHostToDevice, c->stream));
  TRY(launch_psolve(c, c->vh, c->vt));
  TRCHK(c, hipMemcpyAsync(y, c->vt, vbytes(c), hipMemcpyDeviceToHost, c->stream));
  TRCHK(c, hipStreamSynchronize(c->stream));
  return check_gs_error(c);
}

extern "C" int asam_tr_solve(asam_tr_ctx* c, double* x, const double* b, int* maxiter,
                             double* tol, asam_tr_info* info) {
  if (!c || !x || !b || !maxiter || !tol) return trfail(c, ASAM_E_INVALID, "null argument");
  if (!(*tol >= 0.0)) return trfail(c, ASAM_E_INVALID, "tol");
  TRY(check_ready(c, c->ops_ps));
  TRY(check_ready(c, c->ops_mv));
  TRCHK(c, hipSetDevice(c->device));
  TRCHK(c, hipMemcpyAsync(c->vx, x, vbytes(c), hipMemcpyHostToDevice, c->stream));
  TRCHK(c, hipMemcpyAsync(c->vb, b, vbytes(c), hipMemcpyHostToDevice, c->stream));
  TRY(solve_dev(c, c->vx, c->vb, maxiter, tol, info));
  TRCHK(c, hipMemcpyAsync(x, c->vx, vbytes(c), hipMemcpyDeviceToHost, c->stream));
  TRCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int asam_tr_get_tuning(const asam_tr_ctx* c, asam_tr_tuning* t) {
  if (!c || !t) return ASAM_E_INVALID;
  *t = c->tune;
  return 0;
}

extern "C" int asam_tr_set_tuning(asam_tr_ctx* c, const asam_tr_tuning* t) {
  if (!c || !t) return ASAM_E_INVALID;
  if (t->gs_rows < 1 || t->gs_rows > 8) return trfail(c, ASAM_E_INVALID, "gs_rows");
  if (t->gs_chunk != 2 && t->gs_chunk != 4) return trfail(c, ASAM_E_INVALID, "gs_chunk");
  if (t->gs_spin_limit < 0) return trfail(c, ASAM_E_INVALID, "gs_spin_limit");
  if ((2 * (size_t)t->gs_rows * c->d.Z + 2 * (size_t)t->gs_chunk * kBoundStride) *
          sizeof(double) > 60 * 1024)
    return trfail(c, ASAM_E_INVALID, "gs_rows: LDS");
  c->tune = *t;
  return 0;
}

extern "C" int asam_tr_test_hold_boundary(asam_tr_ctx* c, int hold) {
  if (!c) return ASAM_E_INVALID;
  c->hold_boundary = hold != 0;
  return 0;
}

extern "C" int asam_tr_time_kernel(asam_tr_ctx* c, int which, int reps, double* ms) {
  if (!c || !ms || reps < 1 || which < 0 || which > 2) return trfail(c, ASAM_E_INVALID, "argument");
  TRY(check_ready(c, which == 1 ? c->ops_mv : c->ops_ps));
  TRCHK(c, hipSetDevice(c->device));
  hipEvent_t e0, e1;
  TRCHK(c, hipEventCreate(&e0));
  TRCHK(c, hipEventCreate(&e1));
  int rc = 0;
  auto once = [&]() -> int {
    if (which == 0) {
      TRY(launch_sweep(c, 1));
    } else if (which == 1) {
      TRY(launch_matvec(c, c->vp, c->vv, 0));
    } else {
      TRY(launch_psolve(c, c->vp, c->vph));
    }
    return 0;
  };
  if (which == 0) {  // sweeps need a skewed right-hand side
    k_skew_vec<<<blocks_for((long)c->nc * c->d.m, 256), 256, 0, c->stream>>>(c->d, c->nc, c->vp,
                                                                              c->bs);
  }
  rc = once();
  if (!rc) {
    (void)hipEventRecord(e0, c->stream);
    for (int q = 0; q < reps && !rc; ++q) rc = once();
    (void)hipEventRecord(e1, c->stream);
    (void)hipEventSynchronize(e1);
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = t / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  return check_gs_error(c);
}

extern "C" int asam_tr_dev_alloc(asam_tr_ctx* c, void** ptr, size_t bytes) {
  if (!c || !ptr) return ASAM_E_INVALID;
  TRCHK(c, hipSetDevice(c->device));
  TRCHK(c, hipMalloc(ptr, bytes ? bytes : 1));
  return 0;
}
extern "C" int asam_tr_dev_free(asam_tr_ctx* c, void* ptr) {
  if (!c) return ASAM_E_INVALID;
  if (ptr) TRCHK(c, hipFree(ptr));
  return 0;
}
extern "C" int asam_tr_dev_upload(asam_tr_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || !dst || !src) return ASAM_E_INVALID;
  TRCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  TRCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
extern "C" int asam_tr_dev_download(asam_tr_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || !dst || !src) return ASAM_E_INVALID;
  TRCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  TRCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
