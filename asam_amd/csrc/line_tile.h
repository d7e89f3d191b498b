// line_tile.h -- vertical-line preconditioner solve of one 64-column tile (shared by
// the single-domain and the ensemble kernels).
#pragma once
#include <type_traits>

#include "kcommon.h"

namespace asam {

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

// One tile of the line solve: 64 columns (i = 64 bxl + lane) of row j, gxl tiles per
// row.  One thread per vertical column; each sweep works through windows of P levels whose
// loads are all issued before the recurrence consumes them (the recurrence itself is
// sequential, so memory-level parallelism has to come from the window).
template <int NP, int TR, int MODE, int LDS, int P, typename CT, typename VT>
__device__ __forceinline__ void line_tile(const KLine<CT, VT>& a, int bxl, int j, int gxl,
                                          CT* stage, double (&acc)[1]) {
  const int nx = a.nx, nz = a.nz;
  const long sz = (long)nx * a.ny;
  const int tid = threadIdx.x;
  const int i = bxl * 64 + tid;
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
      f0 = (((long)j * gxl + bxl) * nz) * 64 + tid;
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
    // A window holds the loads of P levels.  The sweeps are software-pipelined: the loads
    // of the next window are issued before the recurrence of the current one runs, so
    // two windows are in flight per wavefront.  Each branch below holds a load block and
    // a compute block together, so no control-flow join separates a window's loads from
    // their use (a join would make the compiler drain every outstanding load there).
    struct FWin {
      V r[P], p[P], v[P];
      M l[P], t[P];
    };
    struct BWin {
      V y[P];
      M w[P], t[P];
    };
    V yprev{0, 0};
    // levels past the top of the column (a partial last window) load level nz-1
    auto fload = [&](int kb, FWin& W, auto first_t) {
      constexpr bool FIRST = decltype(first_t)::value;
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int k = kb + q < nz ? kb + q : nz - 1;
        const long c = c0 + sz * k;
        if constexpr (MODE == LINE_PLAIN) W.r[q] = ldv(SRC, c);
        if constexpr (MODE == LINE_BICG_P) {
          W.r[q] = ldv(SRC, c);
          if constexpr (!FIRST) { W.p[q] = ldv((const VT*)PP, c); W.v[q] = ldv(VV, c); }
        }
        if constexpr (MODE == LINE_BICG_S) { W.r[q] = ldv((const VT*)PP, c); W.v[q] = ldv(VV, c); }
        const long f = f0 + fs * k;
        if constexpr (TR == 0) { W.l[q] = ldm(PL, f); W.t[q] = ldm(PT, f); }
        else W.l[q] = ldm(PW, f0 + fs * (k > 0 ? k - 1 : 0));  // unused at k = 0
      }
    };
    auto fcomp = [&](int kb, auto chk, const FWin& W, auto first_t) {
      constexpr bool CHK = decltype(chk)::value;
      constexpr bool FIRST = decltype(first_t)::value;
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int k = kb + q;
        if (!CHK || k < nz) {
          const long c = c0 + sz * k;
          V in;
          if constexpr (MODE == LINE_PLAIN) in = V{W.r[q].a * sc, W.r[q].b * sc};
          else if constexpr (MODE == LINE_BICG_P) {
            if constexpr (FIRST) in = W.r[q];
            else
              in = V{W.r[q].a + beta * (W.p[q].a - omega * W.v[q].a),
                     W.r[q].b + beta * (W.p[q].b - omega * W.v[q].b)};
            stv(PP, c, in);
          } else {
            in = V{W.r[q].a - alpha * W.v[q].a, W.r[q].b - alpha * W.v[q].b};
            stv(PP, c, in);
            acc[0] += (double)in.a * (double)in.a + (double)in.b * (double)in.b;
          }
          V y;
          if constexpr (TR == 0) {
            const V ly = mul(W.l[q], yprev);
            y = mul(W.t[q], V{in.a - ly.a, in.b - ly.b});
          } else {
            const V wy = k > 0 ? mulT(W.l[q], yprev) : V{0, 0};
            y = V{in.a - wy.a, in.b - wy.b};
          }
          stage_put(k, c, y);
          yprev = y;
        }
      }
    };
    // --- backward sweep: z_k = y_k - w_k z_k+1   (transposed: t_k^-T (y_k - l_k+1^T z_k+1))
    V znext{0, 0};
    auto bload = [&](int kb, BWin& W) {
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const int k = kb + q < nz ? kb + q : nz - 1;
        const long c = c0 + sz * k;
        W.y[q] = stage_get(k, c);
        const long f = f0 + fs * k;
        if constexpr (TR == 0) W.w[q] = ldm(PW, f);
        else {
          W.w[q] = ldm(PL, k + 1 < nz ? f + fs : f);  // unused at k = nz - 1
          W.t[q] = ldm(PT, f);
        }
      }
    };
    auto bcomp = [&](int kb, auto chk, const BWin& W) {
      constexpr bool CHK = decltype(chk)::value;
#pragma unroll
      for (int q = P - 1; q >= 0; --q) {
        const int k = kb + q;
        if (!CHK || k < nz) {
          const long c = c0 + sz * k;
          V z;
          if constexpr (TR == 0) {
            const V wz = mul(W.w[q], znext);
            z = V{W.y[q].a - wz.a, W.y[q].b - wz.b};
          } else {
            const V lz = k + 1 < nz ? mulT(W.w[q], znext) : V{0, 0};
            z = mulT(W.t[q], V{W.y[q].a - lz.a, W.y[q].b - lz.b});
          }
          stv(OUT, c, V{-z.a, -z.b});
          znext = z;
        }
      }
    };
    using Full = std::false_type;
    using Chk = std::true_type;
    const int nwin = (nz + P - 1) / P;  // the last window may be partial
    auto sweeps = [&](auto first_t) {
      FWin X, Y;
      fload(0, X, first_t);
      for (int w = 0;; ++w) {  // X holds window w
        if (w + 1 < nwin) {
          fload((w + 1) * P, Y, first_t);
          fcomp(w * P, Full{}, X, first_t);
        } else {
          fcomp(w * P, Chk{}, X, first_t);
          break;
        }
        ++w;  // Y holds window w
        if (w + 1 < nwin) {
          fload((w + 1) * P, X, first_t);
          fcomp(w * P, Full{}, Y, first_t);
        } else {
          fcomp(w * P, Chk{}, Y, first_t);
          break;
        }
      }
    };
    if constexpr (MODE == LINE_BICG_P) {
      if (first) sweeps(std::true_type{}); else sweeps(std::false_type{});
    } else {
      sweeps(std::false_type{});
    }
    {
      BWin X, Y;
      bload((nwin - 1) * P, X);
      if (nwin > 1) {
        bload((nwin - 2) * P, Y);
        bcomp((nwin - 1) * P, Chk{}, X);
        for (int w = nwin - 2;; --w) {  // Y holds window w
          if (w > 0) {
            bload((w - 1) * P, X);
            bcomp(w * P, Full{}, Y);
          } else {
            bcomp(0, Full{}, Y);
            break;
          }
          --w;  // X holds window w
          if (w > 0) {
            bload((w - 1) * P, Y);
            bcomp(w * P, Full{}, X);
          } else {
            bcomp(0, Full{}, X);
            break;
          }
        }
      } else {
        bcomp(0, Chk{}, X);
      }
    }
  }
}

}  // namespace asam
