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
  int kl;  // LDS staging: levels k < kl are staged in LDS, the rest in the output vector
  KTail fold;  // scalar tail folded into the launch (k_line only; kind < 0: none)
};

template <typename T> struct M4 { T m[4]; };

// Factor entries are read once per sweep and by no other kernel.  NT = 1 loads them
// nontemporal, so that factors larger than the Infinity Cache do not push out the vectors
// and operator arrays the neighbouring kernels reuse; NT = 0 lets factors that fit stay
// resident across solves.
template <int NT, typename T>
__device__ __forceinline__ M4<T> ld4(const T* __restrict__ p, long c) {
  M4<T> r;
  if constexpr (sizeof(T) == 4) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    const f4v* q = reinterpret_cast<const f4v*>(p + 4 * c);
    f4v v;
    if constexpr (NT) v = __builtin_nontemporal_load(q); else v = *q;
    r.m[0] = v.x; r.m[1] = v.y; r.m[2] = v.z; r.m[3] = v.w;
  } else {
    typedef double d2v __attribute__((ext_vector_type(2)));
    const d2v* q = reinterpret_cast<const d2v*>(p + 4 * c);
    d2v v0, v1;
    if constexpr (NT) {
      v0 = __builtin_nontemporal_load(q);
      v1 = __builtin_nontemporal_load(q + 1);
    } else {
      v0 = q[0];
      v1 = q[1];
    }
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
//
// NT: cache policy of the factor loads, of the two vectors only this kernel reads (p in
// the p-step, v in the s-step) and of the p-step's store of p (read again only by the next
// iteration's p-step), see ld4.  PBW: levels per window of the backward sweep (0: P); the
// backward sweep runs after the whole forward sweep, its partial window at the top first.
template <int NP, int TR, int MODE, int LDS, int P, typename CT, typename VT, int NT = 1, int PBW = 0>
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
    auto ldvo = [&](const VT* p, long c) -> V {  // a vector this kernel alone reads
      if constexpr (NT == 0) return ldv(p, c);
      else if constexpr (NP == 1) return V{(CT)__builtin_nontemporal_load(p + c), (CT)0};
      else {
        typedef VT v2v __attribute__((ext_vector_type(2)));
        const v2v t = __builtin_nontemporal_load(reinterpret_cast<const v2v*>(p + 2 * c));
        return V{(CT)t.x, (CT)t.y};
      }
    };
    auto stvn = [&](VT* p, long c, V v) {
      if constexpr (NP == 1) __builtin_nontemporal_store((VT)v.a, p + c);
      else {
        typedef VT v2v __attribute__((ext_vector_type(2)));
        v2v t; t.x = (VT)v.a; t.y = (VT)v.b;
        __builtin_nontemporal_store(t, reinterpret_cast<v2v*>(p + 2 * c));
      }
    };
    auto stv = [&](VT* p, long c, V v) {
      if constexpr (NP == 1) p[c] = (VT)v.a; else st2(p, c, v.a, v.b);
    };
    auto ldm = [&](const CT* p, long c) -> M {
      if constexpr (NP == 1) {
        if constexpr (NT) return M{{__builtin_nontemporal_load(p + c), (CT)0, (CT)0, (CT)0}};
        else return M{{p[c], (CT)0, (CT)0, (CT)0}};
      } else {
        return ld4<NT>(p, c);
      }
    };
    // a lower block L: for 2x2 blocks the three words (tz, v0, v1) of
    // L = (tz, 1)^T (v0, v1), in m[0..2]
    auto ldl = [&](const CT* p, long c) -> M {
      if constexpr (NP == 1) {
        return ldm(p, c);
      } else {
        const CT* q = p + 3 * c;
        if constexpr (NT)
          return M{{__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1),
                    __builtin_nontemporal_load(q + 2), (CT)0}};
        else
          return M{{q[0], q[1], q[2], (CT)0}};
      }
    };
    auto mull = [&](const M& l, V y) -> V {  // L y = (tz s, s), s = v0 y.a + v1 y.b
      if constexpr (NP == 1) {
        return V{l.m[0] * y.a, (CT)0};
      } else {
        const CT s = l.m[1] * y.a + l.m[2] * y.b;
        return V{l.m[0] * s, s};
      }
    };
    auto mullT = [&](const M& l, V z) -> V {  // L^T z = (v0 t, v1 t), t = tz z.a + z.b
      if constexpr (NP == 1) {
        return V{l.m[0] * z.a, (CT)0};
      } else {
        const CT t = l.m[0] * z.a + z.b;
        return V{l.m[1] * t, l.m[2] * t};
      }
    };
    auto mul = [&](const M& m, V v) -> V {  // m * v
      if constexpr (NP == 1) return V{m.m[0] * v.a, (CT)0};
      else return mv(m, v);
    };
    auto mulT = [&](const M& m, V v) -> V {  // m^T * v
      if constexpr (NP == 1) return V{m.m[0] * v.a, (CT)0};
      else return mtv(m, v);
    };
    // Forward-sweep results: the lowest kl levels in LDS (written first, read last, so
    // the ones a round trip through memory would most likely miss in the caches), the
    // levels above in the output vector (their backward reads follow the writes closely).
    // k is the same in every lane, so the test is a uniform branch.
    auto stage_put = [&](int k, long c, V y) {
      if constexpr (LDS) {
        if (k < a.kl) {
          stage[(NP * k) * 64 + tid] = y.a;
          if constexpr (NP == 2) stage[(2 * k + 1) * 64 + tid] = y.b;
          return;
        }
      }
      stv(OUT, c, y);
    };
    auto stage_get = [&](int k, long c) -> V {
      if constexpr (LDS) {
        if (k < a.kl) {
          if constexpr (NP == 1) return V{stage[k * 64 + tid], (CT)0};
          else return V{stage[(2 * k) * 64 + tid], stage[(2 * k + 1) * 64 + tid]};
        }
      }
      return ldv((const VT*)OUT, c);
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
            if (!first) { p_[q] = ldvo((const VT*)PP, c); v_[q] = ldv(VV, c); }
          }
          if constexpr (MODE == LINE_BICG_S) { r_[q] = ldv((const VT*)PP, c); v_[q] = ldvo(VV, c); }
          const long f = f0 + fs * k;
          if constexpr (TR == 0) { l_[q] = ldl(PL, f); t_[q] = ldm(PT, f); }
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
            if constexpr (NT) stvn(PP, c, in); else stv(PP, c, in);
          } else {
            in = V{r_[q].a - alpha * v_[q].a, r_[q].b - alpha * v_[q].b};
            stv(PP, c, in);
            acc[0] += (double)in.a * (double)in.a + (double)in.b * (double)in.b;
          }
          V y;
          if constexpr (TR == 0) {
            const V ly = mull(l_[q], yprev);
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
    constexpr int PB = PBW > 0 ? PBW : P;
    auto bwd = [&](int kb, auto chk) {
      constexpr bool CHK = decltype(chk)::value;
      V y_[PB];
      M w_[PB], t_[PB];
#pragma unroll
      for (int q = 0; q < PB; ++q) {
        const int k = kb + q;
        if (!CHK || k < nz) {
          const long c = c0 + sz * k;
          y_[q] = stage_get(k, c);
          const long f = f0 + fs * k;
          if constexpr (TR == 0) w_[q] = ldm(PW, f);
          else {
            w_[q] = (k + 1 < nz) ? ldl(PL, f + fs) : M{{0, 0, 0, 0}};
            t_[q] = ldm(PT, f);
          }
        }
      }
#pragma unroll
      for (int q = PB - 1; q >= 0; --q) {
        const int k = kb + q;
        if (!CHK || k < nz) {
          const long c = c0 + sz * k;
          V z;
          if constexpr (TR == 0) {
            const V wz = mul(w_[q], znext);
            z = V{y_[q].a - wz.a, y_[q].b - wz.b};
          } else {
            const V lz = mullT(w_[q], znext);
            z = mulT(t_[q], V{y_[q].a - lz.a, y_[q].b - lz.b});
          }
          stv(OUT, c, V{-z.a, -z.b});
          znext = z;
        }
      }
    };
    const int nfull = nz / P;
    for (int w = 0; w < nfull; ++w) fwd(w * P, std::false_type{});
    if (nfull * P < nz) fwd(nfull * P, std::true_type{});
    const int nfb = nz / PB;
    if (nfb * PB < nz) bwd(nfb * PB, std::true_type{});
    for (int w = nfb - 1; w >= 0; --w) bwd(w * PB, std::false_type{});
  }
}

}  // namespace asam
