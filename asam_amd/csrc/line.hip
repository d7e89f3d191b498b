// line.hip -- vertical-line preconditioner solve (gfx950): one thread per column, 64
// columns per wavefront (coalesced along x at every level), loads issued a window of
// levels ahead of the Thomas recurrence, forward-sweep results staged in LDS.
#include "kcommon.h"

namespace asam {

// ===================================================================== line solve
template <typename CT, typename VT>
struct KLine {
  Coef<CT> c;
  int nx, ny, nz, tiled;
  const VT* src;
  VT* p;
  const VT* v;
  VT* out;
  const double* scale;
  const KState* st;
  double* partial;
  const int* done;
  int qslot;
};

int line_blocks(const Geom& g) { return ((g.nx + 63) / 64) * g.ny; }

// Dynamic LDS a line-solve workgroup may take: the device's per-workgroup limit
// (capped at 64 KiB) minus 4 KiB kept for the static reduction scratch.
static size_t lds_budget() {
  static int cached = -1;
  if (cached < 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      v = 0;
    if (v > 64 * 1024) v = 64 * 1024;
    cached = v > 4096 ? v - 4096 : 0;
  }
  return (size_t)cached;
}

template <typename T> struct M4 { T m[4]; };

template <typename T>
__device__ __forceinline__ M4<T> ld4(const T* __restrict__ p, long c) {
  M4<T> r;
  if constexpr (sizeof(T) == 4) {
    const float4 v = *reinterpret_cast<const float4*>(p + 4 * c);
    r.m[0] = v.x; r.m[1] = v.y; r.m[2] = v.z; r.m[3] = v.w;
  } else {
    const double2 v0 = *reinterpret_cast<const double2*>(p + 4 * c);
    const double2 v1 = *reinterpret_cast<const double2*>(p + 4 * c + 2);
    r.m[0] = v0.x; r.m[1] = v0.y; r.m[2] = v1.x; r.m[3] = v1.y;
  }
  return r;
}
template <typename T>
__device__ __forceinline__ V2<T> mv(const M4<T>& q, V2<T> v) {  // q * v, column-major 2x2
  return {q.m[0] * v.a + q.m[2] * v.b, q.m[1] * v.a + q.m[3] * v.b};
}
template <typename T>
__device__ __forceinline__ V2<T> mtv(const M4<T>& q, V2<T> v) {  // q^T * v
  return {q.m[0] * v.a + q.m[1] * v.b, q.m[2] * v.a + q.m[3] * v.b};
}

// One thread per vertical column; each sweep works through windows of P levels whose
// loads are all issued before the recurrence consumes them (the recurrence itself is
// sequential, so memory-level parallelism has to come from the window).
template <int NP, int TR, int MODE, int LDS, int P, typename CT, typename VT>
__global__ __launch_bounds__(64) void k_line(KLine<CT, VT> a) {
  if (a.done && *a.done) return;
  extern __shared__ unsigned char lds_raw[];
  CT* stage = reinterpret_cast<CT*>(lds_raw);
  const int nx = a.nx, nz = a.nz;
  const long sz = (long)nx * a.ny;
  const int tid = threadIdx.x;
  const int i = blockIdx.x * 64 + tid, j = blockIdx.y;
  const int blk = blockIdx.x + gridDim.x * blockIdx.y;
  double acc[1] = {0.0};
  const CT* __restrict__ PL = a.c.pl;
  const CT* __restrict__ PT = a.c.pt;
  const CT* __restrict__ PW = a.c.pw;
  const VT* __restrict__ SRC = a.src;
  const VT* __restrict__ VV = a.v;
  VT* __restrict__ OUT = a.out;
  VT* __restrict__ PP = a.p;
  using V = V2<CT>;
  using M = M4<CT>;
  if (i < nx) {
    const long c0 = i + (long)nx * j;
    // factor entries: tiled (64-column tiles, levels contiguous) or x-fastest
    long f0, fs;
    if (a.tiled) {
      f0 = (((long)j * gridDim.x + blockIdx.x) * nz) * 64 + tid;
      fs = 64;
    } else {
      f0 = c0;
      fs = sz;
    }
    CT alpha = 0, beta = 0, omega = 0, sc = 1;
    bool first = false;
    if constexpr (MODE == LINE_BICG_P) {
      first = a.st->iter == 1;
      beta = (CT)a.st->beta; omega = (CT)a.st->omega;
    }
    if constexpr (MODE == LINE_BICG_S) alpha = (CT)a.st->alpha;
    if constexpr (MODE == LINE_PLAIN) { if (a.scale) sc = (CT)(*a.scale); }
    // --- per-level helpers (NP = 1 uses the .a lane of V / m[0] of M only)
    auto ldv = [&](const VT* p, long c) -> V {
      if constexpr (NP == 1) return V{(CT)p[c], (CT)0};
      else { V2<VT> t = ld2(p, c); return V{(CT)t.a, (CT)t.b}; }
    };
    auto stv = [&](VT* p, long c, V v) {
      if constexpr (NP == 1) p[c] = (VT)v.a; else st2(p, c, v.a, v.b);
    };
    auto ldm = [&](const CT* p, long c) -> M {
      if constexpr (NP == 1) return M{{p[c], (CT)0, (CT)0, (CT)0}};
      else return ld4(p, c);
    };
    auto mul = [&](const M& m, V v) -> V {  // m * v
      if constexpr (NP == 1) return V{m.m[0] * v.a, (CT)0};
      else return mv(m, v);
    };
    auto mulT = [&](const M& m, V v) -> V {  // m^T * v
      if constexpr (NP == 1) return V{m.m[0] * v.a, (CT)0};
      else return mtv(m, v);
    };
    auto stage_put = [&](int k, long c, V y) {
      if constexpr (LDS) {
        stage[(NP * k) * 64 + tid] = y.a;
        if constexpr (NP == 2) stage[(2 * k + 1) * 64 + tid] = y.b;
      } else {
        stv(OUT, c, y);
      }
    };
    auto stage_get = [&](int k, long c) -> V {
      if constexpr (LDS) {
        if constexpr (NP == 1) return V{stage[k * 64 + tid], (CT)0};
        else return V{stage[(2 * k) * 64 + tid], stage[(2 * k + 1) * 64 + tid]};
      } else {
        return ldv((const VT*)OUT, c);
      }
    };
    // --- forward sweep: y_k = t_k^-1 (in_k - l_k y_k-1)   (transposed: in_k - w_k-1^T y_k-1)
    V yprev{0, 0};
    auto fwd = [&](int kb, auto chk) {
      constexpr bool CHK = decltype(chk)::value;
      V r_[P], p_[P], v_[P];
      M l_[P], t_[P];
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int k = kb + q;
        if (!CHK || k < nz) {
          const long c = c0 + sz * k;
          if constexpr (MODE == LINE_PLAIN) r_[q] = ldv(SRC, c);
          if constexpr (MODE == LINE_BICG_P) {
            r_[q] = ldv(SRC, c);
            if (!first) { p_[q] = ldv((const VT*)PP, c); v_[q] = ldv(VV, c); }
          }
          if constexpr (MODE == LINE_BICG_S) { r_[q] = ldv((const VT*)PP, c); v_[q] = ldv(VV, c); }
          const long f = f0 + fs * k;
          if constexpr (TR == 0) { l_[q] = ldm(PL, f); t_[q] = ldm(PT, f); }
          else l_[q] = (k > 0) ? ldm(PW, f - fs) : M{{0, 0, 0, 0}};
        }
      }
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int k = kb + q;
        if (!CHK || k < nz) {
          const long c = c0 + sz * k;
          V in;
          if constexpr (MODE == LINE_PLAIN) in = V{r_[q].a * sc, r_[q].b * sc};
          else if constexpr (MODE == LINE_BICG_P) {
            in = first ? r_[q]
                       : V{r_[q].a + beta * (p_[q].a - omega * v_[q].a),
                           r_[q].b + beta * (p_[q].b - omega * v_[q].b)};
            stv(PP, c, in);
          } else {
            in = V{r_[q].a - alpha * v_[q].a, r_[q].b - alpha * v_[q].b};
            stv(PP, c, in);
            acc[0] += (double)in.a * (double)in.a + (double)in.b * (double)in.b;
          }
          V y;
          if constexpr (TR == 0) {
            const V ly = mul(l_[q], yprev);
            y = mul(t_[q], V{in.a - ly.a, in.b - ly.b});
          } else {
            const V wy = mulT(l_[q], yprev);
            y = V{in.a - wy.a, in.b - wy.b};
          }
          stage_put(k, c, y);
          yprev = y;
        }
      }
    };
    // --- backward sweep: z_k = y_k - w_k z_k+1   (transposed: t_k^-T (y_k - l_k+1^T z_k+1))
    V znext{0, 0};
    auto bwd = [&](int kb, auto chk) {
      constexpr bool CHK = decltype(chk)::value;
      V y_[P];
      M w_[P], t_[P];
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int k = kb + q;
        if (!CHK || k < nz) {
          const long c = c0 + sz * k;
          y_[q] = stage_get(k, c);
          const long f = f0 + fs * k;
          if constexpr (TR == 0) w_[q] = ldm(PW, f);
          else {
            w_[q] = (k + 1 < nz) ? ldm(PL, f + fs) : M{{0, 0, 0, 0}};
            t_[q] = ldm(PT, f);
          }
        }
      }
#pragma unroll
      for (int q = P - 1; q >= 0; --q) {
        const int k = kb + q;
        if (!CHK || k < nz) {
          const long c = c0 + sz * k;
          V z;
          if constexpr (TR == 0) {
            const V wz = mul(w_[q], znext);
            z = V{y_[q].a - wz.a, y_[q].b - wz.b};
          } else {
            const V lz = mulT(w_[q], znext);
            z = mulT(t_[q], V{y_[q].a - lz.a, y_[q].b - lz.b});
          }
          stv(OUT, c, V{-z.a, -z.b});
          znext = z;
        }
      }
    };
    const int nfull = nz / P;
    for (int w = 0; w < nfull; ++w) fwd(w * P, std::false_type{});
    if (nfull * P < nz) {
      fwd(nfull * P, std::true_type{});
      bwd(nfull * P, std::true_type{});
    }
    for (int w = nfull - 1; w >= 0; --w) bwd(w * P, std::false_type{});
  }
  if constexpr (MODE == LINE_BICG_S) {
    block_reduce_store<1>(acc, a.partial + (long)a.qslot * kRedBlocks, blk, tid, 64);
  }
}

template <int NP, int TR, int LDS, int P, typename CT, typename VT>
static void line_launch(hipStream_t s, const Geom& g, const KLine<CT, VT>& a, int mode,
                        size_t lds) {
  dim3 grid((g.nx + 63) / 64, g.ny);
  switch (mode) {
    case LINE_PLAIN:
      hipLaunchKernelGGL((k_line<NP, TR, LINE_PLAIN, LDS, P, CT, VT>), grid, dim3(64), lds, s, a);
      break;
    case LINE_BICG_P:
      hipLaunchKernelGGL((k_line<NP, TR, LINE_BICG_P, LDS, P, CT, VT>), grid, dim3(64), lds, s, a);
      break;
    default:
      hipLaunchKernelGGL((k_line<NP, TR, LINE_BICG_S, LDS, P, CT, VT>), grid, dim3(64), lds, s, a);
      break;
  }
}

template <int NP, int TR, typename CT, typename VT>
static void line_t(hipStream_t s, const Geom& g, const Coef<CT>& c, const LineReq& r) {
  KLine<CT, VT> a;
  a.c = c; a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.tiled = g.tune.line_tiled;
  a.src = (const VT*)r.src; a.p = (VT*)r.p; a.v = (const VT*)r.v; a.out = (VT*)r.out;
  a.scale = r.scale; a.st = r.st; a.partial = r.partial; a.done = r.done;
  a.qslot = r.qslot;
  const size_t lds = (size_t)g.nz * 64 * NP * sizeof(CT);
  const bool use_lds = r.use_lds && g.tune.line_lds && lds <= lds_budget();
  // default window: 8 levels, 4 for FP64 2x2 blocks (register budget)
  const int P = g.tune.line_window > 0 ? g.tune.line_window
                                       : ((sizeof(CT) == 8 && NP == 2) ? 4 : 8);
  if (use_lds) {
    if (P >= 8) line_launch<NP, TR, 1, 8, CT, VT>(s, g, a, r.mode, lds);
    else if (P >= 4) line_launch<NP, TR, 1, 4, CT, VT>(s, g, a, r.mode, lds);
    else line_launch<NP, TR, 1, 1, CT, VT>(s, g, a, r.mode, lds);
  } else {
    if (P >= 8) line_launch<NP, TR, 0, 8, CT, VT>(s, g, a, r.mode, 0);
    else if (P >= 4) line_launch<NP, TR, 0, 4, CT, VT>(s, g, a, r.mode, 0);
    else line_launch<NP, TR, 0, 1, CT, VT>(s, g, a, r.mode, 0);
  }
}

template <int NP, int TR>
static void line_np(hipStream_t s, const Geom& g, const Coef<double>& c64,
                    const Coef<float>& c32, const LineReq& r) {
  if (r.prec == 0) line_t<NP, TR, double, double>(s, g, c64, r);
  else line_t<NP, TR, float, float>(s, g, c32, r);
}

void launch_line(hipStream_t s, const Geom& g, const Coef<double>& c64, const Coef<float>& c32,
                 const LineReq& r) {
  if (g.np1 == 1) {
    if (r.tr) line_np<1, 1>(s, g, c64, c32, r); else line_np<1, 0>(s, g, c64, c32, r);
  } else {
    if (r.tr) line_np<2, 1>(s, g, c64, c32, r); else line_np<2, 0>(s, g, c64, c32, r);
  }
}

}  // namespace asam
