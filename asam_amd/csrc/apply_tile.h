// apply_tile.h -- operator apply of one tile (shared by the single-domain and the
// ensemble kernels).
//
// Operator (one block, Neumann walls), per cell c with faces f to neighbours n:
//   MultiTriT : (A x)_c = fa V_c x_c + sum_f k_f (DUT_c x_c - DUT_n x_n)
//   MultiTriTB: row1 = fa V x1 + sum_f th_f k'_f (q_c - q_n) + vertical gravity * th_f
//               row2 =    V x2 + sum_f      k'_f (q_c - q_n) + vertical gravity
//               q = DUT x1; gravity flux of the z-face (k,k+1):
//               B = g~ (V_k x2_k + V_k+1 x2_k+1), -B into cell k, +B into cell k+1.
// This is the reference's nested tri-diagonal matrix (ComputeFull/Dual_1D..3D,
// TriDiagT_Mod.f90:32-900; Compute_1D..3D, TriDiagTB_Mod.f90:32-780) written per face:
// u = -k DUT_n, l = -k DUT_c, d += k DUT_c, and the 2x2 gravity entries
// (TriDiagTB_Mod.f90:56-81) with g~ = Grav (beta0 dtP)^2 F / (V_k + V_k+1 + Eps).
#pragma once
#include "kcommon.h"

namespace asam {

// ===================================================================== apply
// Fused options of one apply launch (compile-time, so no branch separates the loads
// of a cell): residual form y = b - Op(x) (AF_B), copy of y (AF_Y2), x_eff = x + dx
// written to xw (AF_DX), dots y.y (AF_YY) and y.u (AF_YU); AF_X0: x is zero, so the
// residual form is y = b outside the dead cells (no x or coefficient loads).
enum { AF_B = 1, AF_Y2 = 2, AF_DX = 4, AF_YY = 8, AF_YU = 16, AF_X0 = 32 };

template <typename CT, typename VX, typename VO, typename VU>
struct KApply {
  Coef<CT> c;
  int nx, ny, nz;
  int gx, gy, kp, ntiles;  // tiles: bx x by columns x kp planes, ntiles in logical order
  int bxw, byw, xcd;
  int band_rows;           // band-ordered variant: rows per band (gy = row tiles per band)
  CT fa;
  const VX* x;
  const float* dx;
  double* xw;
  const double* b;
  VO* y;
  VO* y2;
  const VU* u;
  double* partial;
  const int* done;
  const unsigned char* dead;  // AF_X0: cells whose residual is 0
  KTail tail;                 // fused scalar tail of the dots (kind < 0: none)
};

// Loads of arrays each cell reads exactly once (no neighbour reuse): nontemporal, so they
// do not displace the reused x/dut/face planes from the caches.
template <typename T>
__device__ __forceinline__ T ld_once(const T* p, long m) {
  return __builtin_nontemporal_load(p + m);
}

// One tile of the z-march apply: a thread owns one column (i, j) of tile L and walks
// its planes upwards, carrying the cells below/at/above and the z-face coefficients in
// registers, so each level loads only the next cell up and the four horizontal
// neighbours.  Dots of the fused options accumulate into acc.
template <int NP, int TR, int FL, typename CT, typename VX, typename VO, typename VU>
__device__ __forceinline__ void march_tile(const KApply<CT, VX, VO, VU>& a, int L,
                                           double (&acc)[2]) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const long sy = nx, sz = (long)nx * ny;
  const CT* __restrict__ kx = a.c.kx;
  const CT* __restrict__ ky = a.c.ky;
  const CT* __restrict__ kz = a.c.kz;
  const CT* __restrict__ tx = a.c.tx;
  const CT* __restrict__ ty = a.c.ty;
  const CT* __restrict__ tz = a.c.tz;
  const CT* __restrict__ gzz = a.c.gz;
  const CT* __restrict__ vol = a.c.vol;
  const CT* __restrict__ dut = a.c.dut;
  const VX* __restrict__ X = a.x;
  const float* __restrict__ DX = a.dx;
  // Outputs and the dot operand through restrict copies: the compiler may then issue the
  // loads of a level (u and b included) before the stores of the previous one.
  VO* __restrict__ Y = a.y;
  VO* __restrict__ Y2 = a.y2;
  const VU* __restrict__ U = a.u;
  double* __restrict__ XW = a.xw;
  const double* __restrict__ Bv = a.b;
  const CT fa = a.fa;
  if (L < a.ntiles) {
    const int bx = L % a.gx, rest = L / a.gx, by = rest % a.gy, kt = rest / a.gy;
    const int i = bx * a.bxw + threadIdx.x, j = by * a.byw + threadIdx.y;
    const int k0 = kt * a.kp, k1 = min(nz, k0 + a.kp);
    if (i < nx && j < ny) {
      const long cb = i + sy * j;
      if constexpr ((FL & AF_X0) != 0) {
        // x = 0: every stencil term is a signed zero, so b + s equals b (a zero b may
        // change sign); the walk and the dot accumulation order are those of the full
        // apply below, so the partial sums are the same bits
        static_assert((FL & AF_B) != 0 && (FL & (AF_DX | AF_YU)) == 0, "x = 0 residual form");
        const unsigned char* __restrict__ D = a.dead;
        long c = cb + sz * k0;
        for (int k = k0; k < k1; ++k, c += sz) {
          const bool dead = D[c] != 0;
          if constexpr (NP == 1) {
            const double yv = dead ? 0.0 : Bv[c];
            Y[c] = (VO)yv;
            if constexpr ((FL & AF_Y2) != 0) Y2[c] = (VO)yv;
            if constexpr ((FL & AF_YY) != 0) acc[0] += yv * yv;
          } else {
            const double2 b0 = *reinterpret_cast<const double2*>(Bv + 2 * c);
            const double yv1 = dead ? 0.0 : b0.x, yv2 = dead ? 0.0 : b0.y;
            st2(Y, c, yv1, yv2);
            if constexpr ((FL & AF_Y2) != 0) st2(Y2, c, yv1, yv2);
            if constexpr ((FL & AF_YY) != 0) acc[0] += yv1 * yv1 + yv2 * yv2;
          }
        }
        return;
      }
      const long dW = i > 0 ? -1 : 0, dE = i + 1 < nx ? 1 : 0;
      const long dS = j > 0 ? -sy : 0, dN = j + 1 < ny ? sy : 0;
      if constexpr (NP == 1) {
        auto LX = [&](long m) -> CT {
          CT v = (CT)X[m];
          if constexpr ((FL & AF_DX) != 0) v += (CT)DX[m];
          return v;
        };
        long c = cb + sz * k0;
        CT xm = 0, dm = 0, kzm = 0;
        if (k0 > 0) { xm = LX(c - sz); dm = dut[c - sz]; kzm = kz[c - sz]; }
        CT x0 = LX(c), d0 = dut[c];
        for (int k = k0; k < k1; ++k, c += sz) {
          const long mU = k + 1 < nz ? c + sz : c;
          const CT xp = LX(mU), dp = dut[mU], kz0 = kz[c];
          const long mW = c + dW, mE = c + dE, mS = c + dS, mN = c + dN;
          VU u0 = 0;
          if constexpr ((FL & AF_YU) != 0) u0 = U[c];
          double b0 = 0.0;
          if constexpr ((FL & AF_B) != 0) b0 = Bv[c];
          CT s = fa * vol[c] * x0;
          if constexpr (TR == 0) {
            const CT q0 = d0 * x0;
            s += kx[mW] * (q0 - dut[mW] * LX(mW));
            s += kx[c] * (q0 - dut[mE] * LX(mE));
            s += ky[mS] * (q0 - dut[mS] * LX(mS));
            s += ky[c] * (q0 - dut[mN] * LX(mN));
            s += kzm * (q0 - dm * xm);
            s += kz0 * (q0 - dp * xp);
          } else {
            CT h = kx[mW] * (x0 - LX(mW));
            h += kx[c] * (x0 - LX(mE));
            h += ky[mS] * (x0 - LX(mS));
            h += ky[c] * (x0 - LX(mN));
            h += kzm * (x0 - xm);
            h += kz0 * (x0 - xp);
            s += d0 * h;
          }
          double yv;
          if constexpr ((FL & AF_B) != 0) {
            const bool dead = vol[c] == (CT)0 && kx[mW] + kx[c] + ky[mS] + ky[c] + kzm + kz0 ==
                                                     (CT)0;
            yv = dead ? 0.0 : b0 + (double)s;
          } else {
            yv = -(double)s;
          }
          if constexpr ((FL & AF_DX) != 0) XW[c] = (double)x0;
          Y[c] = (VO)yv;
          if constexpr ((FL & AF_Y2) != 0) Y2[c] = (VO)yv;
          if constexpr ((FL & AF_YY) != 0) acc[0] += yv * yv;
          if constexpr ((FL & AF_YU) != 0) acc[1] += yv * (double)u0;
          xm = x0; dm = d0; kzm = kz0; x0 = xp; d0 = dp;
        }
      } else {
        auto LX = [&](long m) -> V2<CT> {
          V2<VX> v = ld2(X, m);
          V2<CT> r{(CT)v.a, (CT)v.b};
          if constexpr ((FL & AF_DX) != 0) {
            V2<float> e = ld2(DX, m);
            r.a += (CT)e.a; r.b += (CT)e.b;
          }
          return r;
        };
        long c = cb + sz * k0;
        V2<CT> xm{0, 0};
        CT dm = 0, vm = 0, kzm = 0, tzm = 0, gzm = 0;
        if (k0 > 0) {
          const long m = c - sz;
          xm = LX(m); dm = dut[m]; vm = ld_once(vol, m); kzm = ld_once(kz, m);
          tzm = ld_once(tz, m); gzm = ld_once(gzz, m);
        }
        V2<CT> x0 = LX(c);
        CT d0 = dut[c], v0 = ld_once(vol, c);
        for (int k = k0; k < k1; ++k, c += sz) {
          const long mU = k + 1 < nz ? c + sz : c;
          const V2<CT> xp = LX(mU);
          const CT dp = dut[mU], vp = ld_once(vol, mU);
          const CT kz0 = ld_once(kz, c), tz0 = ld_once(tz, c), gz0 = ld_once(gzz, c);
          const long mW = c + dW, mE = c + dE, mS = c + dS, mN = c + dN;
          const V2<CT> xW = LX(mW), xE = LX(mE), xS = LX(mS), xN = LX(mN);
          V2<VU> u0{0, 0};
          if constexpr ((FL & AF_YU) != 0) u0 = ld2(U, c);
          double2 b0 = make_double2(0.0, 0.0);
          if constexpr ((FL & AF_B) != 0) b0 = *reinterpret_cast<const double2*>(Bv + 2 * c);
          CT s1 = fa * v0 * x0.a, s2 = v0 * x0.b;
          if constexpr (TR == 0) {
            const CT q0 = d0 * x0.a;
            auto hf = [&](CT kk, CT th, CT dn, CT xn) {
              const CT t = kk * (q0 - dn * xn);
              s1 += th * t; s2 += t;
            };
            hf(kx[mW], tx[mW], dut[mW], xW.a);
            hf(kx[c], tx[c], dut[mE], xE.a);
            hf(ky[mS], ty[mS], dut[mS], xS.a);
            hf(ky[c], ty[c], dut[mN], xN.a);
            {
              const CT t = kzm * (q0 - dm * xm.a);
              const CT B = gzm * (vm * xm.b + v0 * x0.b);
              s1 += tzm * (t + B); s2 += t + B;
            }
            {
              const CT t = kz0 * (q0 - dp * xp.a);
              const CT B = gz0 * (v0 * x0.b + vp * xp.b);
              s1 += tz0 * (t - B); s2 += t - B;
            }
          } else {
            CT h = 0;
            auto hf = [&](CT kk, CT th, V2<CT> xn) {
              h += kk * ((th * x0.a + x0.b) - (th * xn.a + xn.b));
            };
            hf(kx[mW], tx[mW], xW);
            hf(kx[c], tx[c], xE);
            hf(ky[mS], ty[mS], xS);
            hf(ky[c], ty[c], xN);
            {
              const CT zc = tzm * x0.a + x0.b, zn = tzm * xm.a + xm.b;
              h += kzm * (zc - zn);
              s2 += gzm * v0 * (zc - zn);
            }
            {
              const CT zc = tz0 * x0.a + x0.b, zn = tz0 * xp.a + xp.b;
              h += kz0 * (zc - zn);
              s2 += gz0 * v0 * (zn - zc);
            }
            s1 += d0 * h;
          }
          double yv1, yv2;
          if constexpr ((FL & AF_B) != 0) {
            const bool dead = v0 == (CT)0 && kx[mW] + kx[c] + ky[mS] + ky[c] + kzm + kz0 == (CT)0;
            yv1 = dead ? 0.0 : b0.x + (double)s1;
            yv2 = dead ? 0.0 : b0.y + (double)s2;
          } else {
            yv1 = -(double)s1; yv2 = -(double)s2;
          }
          if constexpr ((FL & AF_DX) != 0) st2(XW, c, (double)x0.a, (double)x0.b);
          st2(Y, c, yv1, yv2);
          if constexpr ((FL & AF_Y2) != 0) st2(Y2, c, yv1, yv2);
          if constexpr ((FL & AF_YY) != 0) acc[0] += yv1 * yv1 + yv2 * yv2;
          if constexpr ((FL & AF_YU) != 0) acc[1] += yv1 * (double)u0.a + yv2 * (double)u0.b;
          xm = x0; dm = d0; vm = v0; kzm = kz0; tzm = tz0; gzm = gz0;
          x0 = xp; d0 = dp; v0 = vp;
        }
      }
    }
  }
}

// Row-blocked z-march: a thread owns rows j0 .. j0+R-1 of column i of tile L (the tile
// spans bxw columns by byw*R rows) and walks its planes upwards with all R cells of a level
// side by side.  A row's south and north neighbours (x, DUT) and its south-face coefficients
// are the adjacent rows' registers at the same level; only the rows just below and above the
// block are loaded.  Rows past ny are clamped to row ny-1 (equal to their neighbour, so a
// wall face difference is exactly zero, as in march_tile) and store nothing.  The arithmetic
// of a cell is march_tile's, term for term; the dots accumulate level by level, rows in order.
// The x and DUT of the level above at the west and east neighbours and in the rows just
// below and above the block are loaded in the iteration that loads the cell above (the
// owners of those cells load them then too) and carried one level, so they are not
// re-read from the caches an iteration later.
// ZP (2x2 blocks): the z-face coefficients and the volume, which no neighbour reads, come
// from one nontemporal 16-byte load per cell of their packed copy (Coef::pkz), loaded with
// the level above and carried: four streams in one.
template <int NP, int TR, int FL, int R, typename CT, typename VX, typename VO, typename VU,
          int ZP = 0>
__device__ __forceinline__ void march_rows(const KApply<CT, VX, VO, VU>& a, int L,
                                           double (&acc)[2]) {
  // multiply-adds fused within a statement only: the bits follow the source, so
  // march_rows_deep (the same statements with another order of loads) gives the same
  // values (hipcc's default also fuses across statements, which depends on code shape)
#pragma clang fp contract(on)
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const long sy = nx, sz = (long)nx * ny;
  const CT* __restrict__ kx = a.c.kx;
  const CT* __restrict__ ky = a.c.ky;
  const CT* __restrict__ kz = a.c.kz;
  const CT* __restrict__ tx = a.c.tx;
  const CT* __restrict__ ty = a.c.ty;
  const CT* __restrict__ tz = a.c.tz;
  const CT* __restrict__ gzz = a.c.gz;
  const CT* __restrict__ vol = a.c.vol;
  const CT* __restrict__ dut = a.c.dut;
  const VX* __restrict__ X = a.x;
  const float* __restrict__ DX = a.dx;
  VO* __restrict__ Y = a.y;
  VO* __restrict__ Y2 = a.y2;
  const VU* __restrict__ U = a.u;
  double* __restrict__ XW = a.xw;
  const double* __restrict__ Bv = a.b;
  const CT fa = a.fa;
  if (L >= a.ntiles) return;
  const int bx = L % a.gx, rest = L / a.gx, by = rest % a.gy, kt = rest / a.gy;
  const int i = bx * a.bxw + threadIdx.x;
  const int j0 = (by * a.byw + threadIdx.y) * R;
  if (i >= nx || j0 >= ny) return;
  const int k0 = kt * a.kp, k1 = min(nz, k0 + a.kp);
  long cb[R];  // cell index of each row at level 0 (rows past ny: row ny-1)
  bool ok[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    ok[r] = j0 + r < ny;
    cb[r] = i + sy * (ok[r] ? j0 + r : ny - 1);
  }
  // the rows just below and above the block (a wall: the block's own edge row)
  const long cbS = j0 > 0 ? cb[0] - sy : cb[0];
  const long cbN = j0 + R < ny ? cb[R - 1] + sy : cb[R - 1];
  if constexpr ((FL & AF_X0) != 0) {
    static_assert((FL & AF_B) != 0 && (FL & (AF_DX | AF_YU)) == 0, "x = 0 residual form");
    const unsigned char* __restrict__ D = a.dead;
    for (int k = k0; k < k1; ++k) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const long c = cb[r] + sz * k;
        const bool dead = D[c] != 0;
        if constexpr (NP == 1) {
          const double yv = dead ? 0.0 : Bv[c];
          if (ok[r]) {
            Y[c] = (VO)yv;
            if constexpr ((FL & AF_Y2) != 0) Y2[c] = (VO)yv;
            if constexpr ((FL & AF_YY) != 0) acc[0] += yv * yv;
          }
        } else {
          const double2 b0 = *reinterpret_cast<const double2*>(Bv + 2 * c);
          const double yv1 = dead ? 0.0 : b0.x, yv2 = dead ? 0.0 : b0.y;
          if (ok[r]) {
            st2(Y, c, yv1, yv2);
            if constexpr ((FL & AF_Y2) != 0) st2(Y2, c, yv1, yv2);
            if constexpr ((FL & AF_YY) != 0) acc[0] += yv1 * yv1 + yv2 * yv2;
          }
        }
      }
    }
    return;
  }
  const long dW = i > 0 ? -1 : 0, dE = i + 1 < nx ? 1 : 0;
  if constexpr (NP == 1) {
    auto LX = [&](long m) -> CT {
      CT v = (CT)X[m];
      if constexpr ((FL & AF_DX) != 0) v += (CT)DX[m];
      return v;
    };
    CT xm[R], dm[R], kzm[R], x0[R], d0[R], xw0[R], xe0[R], dw0[R], de0[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const long c = cb[r] + sz * k0;
      xm[r] = 0; dm[r] = 0; kzm[r] = 0;
      if (k0 > 0) { xm[r] = LX(c - sz); dm[r] = dut[c - sz]; kzm[r] = kz[c - sz]; }
      x0[r] = LX(c); d0[r] = dut[c];
      xw0[r] = LX(c + dW); xe0[r] = LX(c + dE); dw0[r] = 0; de0[r] = 0;
      if constexpr (TR == 0) { dw0[r] = dut[c + dW]; de0[r] = dut[c + dE]; }
    }
    CT xS0 = LX(cbS + sz * k0), dS0 = 0, xN1 = LX(cbN + sz * k0), dN1 = 0;
    if constexpr (TR == 0) { dS0 = dut[cbS + sz * k0]; dN1 = dut[cbN + sz * k0]; }
    for (int k = k0; k < k1; ++k) {
      const long off = sz * k;
      const bool up = k + 1 < nz;
      const long offU = up ? off + sz : off;
      const CT xSp = LX(cbS + offU), xNp = LX(cbN + offU);
      CT dSp = 0, dNp = 0;
      if constexpr (TR == 0) { dSp = dut[cbS + offU]; dNp = dut[cbN + offU]; }
      CT xp[R], dp[R], kz0[R], kyc[R], xwp[R], xep[R], dwp[R], dep[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const long c = cb[r] + off, mU = up ? c + sz : c;
        xp[r] = LX(mU); dp[r] = dut[mU]; kz0[r] = kz[c]; kyc[r] = ky[c];
        xwp[r] = LX(mU + dW); xep[r] = LX(mU + dE); dwp[r] = 0; dep[r] = 0;
        if constexpr (TR == 0) { dwp[r] = dut[mU + dW]; dep[r] = dut[mU + dE]; }
      }
      const CT kyS0 = ky[cbS + off];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const long c = cb[r] + off;
        const long mW = c + dW, mE = c + dE;
        const CT xs = r > 0 ? x0[r - 1] : xS0, ds = r > 0 ? d0[r - 1] : dS0;
        const CT xn = r + 1 < R ? x0[r + 1] : xN1, dn = r + 1 < R ? d0[r + 1] : dN1;
        const CT kys = r > 0 ? kyc[r - 1] : kyS0;
        VU u0 = 0;
        if constexpr ((FL & AF_YU) != 0) u0 = U[c];
        double b0 = 0.0;
        if constexpr ((FL & AF_B) != 0) b0 = Bv[c];
        CT s = fa * vol[c] * x0[r];
        if constexpr (TR == 0) {
          const CT q0 = d0[r] * x0[r];
          s += kx[mW] * (q0 - dw0[r] * xw0[r]);
          s += kx[c] * (q0 - de0[r] * xe0[r]);
          s += kys * (q0 - ds * xs);
          s += kyc[r] * (q0 - dn * xn);
          s += kzm[r] * (q0 - dm[r] * xm[r]);
          s += kz0[r] * (q0 - dp[r] * xp[r]);
        } else {
          CT h = kx[mW] * (x0[r] - xw0[r]);
          h += kx[c] * (x0[r] - xe0[r]);
          h += kys * (x0[r] - xs);
          h += kyc[r] * (x0[r] - xn);
          h += kzm[r] * (x0[r] - xm[r]);
          h += kz0[r] * (x0[r] - xp[r]);
          s += d0[r] * h;
        }
        double yv;
        if constexpr ((FL & AF_B) != 0) {
          const bool dead = vol[c] == (CT)0 &&
                            kx[mW] + kx[c] + kys + kyc[r] + kzm[r] + kz0[r] == (CT)0;
          yv = dead ? 0.0 : b0 + (double)s;
        } else {
          yv = -(double)s;
        }
        if (ok[r]) {
          if constexpr ((FL & AF_DX) != 0) XW[c] = (double)x0[r];
          Y[c] = (VO)yv;
          if constexpr ((FL & AF_Y2) != 0) Y2[c] = (VO)yv;
          if constexpr ((FL & AF_YY) != 0) acc[0] += yv * yv;
          if constexpr ((FL & AF_YU) != 0) acc[1] += yv * (double)u0;
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        xm[r] = x0[r]; dm[r] = d0[r]; kzm[r] = kz0[r]; x0[r] = xp[r]; d0[r] = dp[r];
        xw0[r] = xwp[r]; xe0[r] = xep[r]; dw0[r] = dwp[r]; de0[r] = dep[r];
      }
      xS0 = xSp; dS0 = dSp; xN1 = xNp; dN1 = dNp;
    }
  } else {
    auto LX = [&](long m) -> V2<CT> {
      V2<VX> v = ld2(X, m);
      V2<CT> r{(CT)v.a, (CT)v.b};
      if constexpr ((FL & AF_DX) != 0) {
        V2<float> e = ld2(DX, m);
        r.a += (CT)e.a; r.b += (CT)e.b;
      }
      return r;
    };
    V2<CT> xm[R], x0[R], xw0[R], xe0[R];
    CT dm[R], vm[R], kzm[R], tzm[R], gzm[R], d0[R], v0[R], dw0[R], de0[R];
    // ZP: the z-face coefficients of the current level, loaded with the level's volume
    CT kzc[R], tzc[R], gzc[R];
    typedef CT v4 __attribute__((ext_vector_type(4)));
    const v4* __restrict__ PKZ = reinterpret_cast<const v4*>(a.c.pkz);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const long c = cb[r] + sz * k0;
      xm[r] = V2<CT>{0, 0}; dm[r] = 0; vm[r] = 0; kzm[r] = 0; tzm[r] = 0; gzm[r] = 0;
      if (k0 > 0) {
        const long m = c - sz;
        xm[r] = LX(m); dm[r] = dut[m];
        if constexpr (ZP) {
          const v4 q = __builtin_nontemporal_load(PKZ + m);
          kzm[r] = q.x; tzm[r] = q.y; gzm[r] = q.z; vm[r] = q.w;
        } else {
          vm[r] = ld_once(vol, m); kzm[r] = ld_once(kz, m);
          tzm[r] = ld_once(tz, m); gzm[r] = ld_once(gzz, m);
        }
      }
      x0[r] = LX(c); d0[r] = dut[c];
      if constexpr (ZP) {
        const v4 q = __builtin_nontemporal_load(PKZ + c);
        kzc[r] = q.x; tzc[r] = q.y; gzc[r] = q.z; v0[r] = q.w;
      } else {
        v0[r] = ld_once(vol, c);
      }
      xw0[r] = LX(c + dW); xe0[r] = LX(c + dE); dw0[r] = 0; de0[r] = 0;
      if constexpr (TR == 0) { dw0[r] = dut[c + dW]; de0[r] = dut[c + dE]; }
    }
    V2<CT> xS0 = LX(cbS + sz * k0), xN1 = LX(cbN + sz * k0);
    CT dS0 = 0, dN1 = 0;
    if constexpr (TR == 0) { dS0 = dut[cbS + sz * k0]; dN1 = dut[cbN + sz * k0]; }
    for (int k = k0; k < k1; ++k) {
      const long off = sz * k;
      const bool up = k + 1 < nz;
      const long offU = up ? off + sz : off;
      const V2<CT> xSp = LX(cbS + offU), xNp = LX(cbN + offU);
      CT dSp = 0, dNp = 0;
      if constexpr (TR == 0) { dSp = dut[cbS + offU]; dNp = dut[cbN + offU]; }
      V2<CT> xp[R], xwp[R], xep[R];
      CT dp[R], vp[R], kz0[R], tz0[R], gz0[R], kyc[R], tyc[R], dwp[R], dep[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const long c = cb[r] + off, mU = up ? c + sz : c;
        xp[r] = LX(mU); dp[r] = dut[mU];
        if constexpr (ZP) {
          const v4 q = __builtin_nontemporal_load(PKZ + mU);
          vp[r] = q.w;
          kz0[r] = kzc[r]; tz0[r] = tzc[r]; gz0[r] = gzc[r];
          kzc[r] = q.x; tzc[r] = q.y; gzc[r] = q.z;
        } else {
          vp[r] = ld_once(vol, mU);
          kz0[r] = ld_once(kz, c); tz0[r] = ld_once(tz, c); gz0[r] = ld_once(gzz, c);
        }
        kyc[r] = ky[c]; tyc[r] = ty[c];
        xwp[r] = LX(mU + dW); xep[r] = LX(mU + dE); dwp[r] = 0; dep[r] = 0;
        if constexpr (TR == 0) { dwp[r] = dut[mU + dW]; dep[r] = dut[mU + dE]; }
      }
      const long cS = cbS + off;
      const CT kyS0 = ky[cS], tyS0 = ty[cS];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const long c = cb[r] + off;
        const long mW = c + dW, mE = c + dE;
        const V2<CT> xW = xw0[r], xE = xe0[r];
        const V2<CT> xs = r > 0 ? x0[r - 1] : xS0, xn = r + 1 < R ? x0[r + 1] : xN1;
        const CT ds = r > 0 ? d0[r - 1] : dS0, dn = r + 1 < R ? d0[r + 1] : dN1;
        const CT kys = r > 0 ? kyc[r - 1] : kyS0, tys = r > 0 ? tyc[r - 1] : tyS0;
        V2<VU> u0{0, 0};
        if constexpr ((FL & AF_YU) != 0) u0 = ld2(U, c);
        double2 b0 = make_double2(0.0, 0.0);
        if constexpr ((FL & AF_B) != 0) b0 = *reinterpret_cast<const double2*>(Bv + 2 * c);
        CT s1 = fa * v0[r] * x0[r].a, s2 = v0[r] * x0[r].b;
        if constexpr (TR == 0) {
          const CT q0 = d0[r] * x0[r].a;
          auto hf = [&](CT kk, CT th, CT dnb, CT xnb) {
            const CT t = kk * (q0 - dnb * xnb);
            s1 += th * t; s2 += t;
          };
          hf(kx[mW], tx[mW], dw0[r], xW.a);
          hf(kx[c], tx[c], de0[r], xE.a);
          hf(kys, tys, ds, xs.a);
          hf(kyc[r], tyc[r], dn, xn.a);
          {
            const CT t = kzm[r] * (q0 - dm[r] * xm[r].a);
            const CT B = gzm[r] * (vm[r] * xm[r].b + v0[r] * x0[r].b);
            s1 += tzm[r] * (t + B); s2 += t + B;
          }
          {
            const CT t = kz0[r] * (q0 - dp[r] * xp[r].a);
            const CT B = gz0[r] * (v0[r] * x0[r].b + vp[r] * xp[r].b);
            s1 += tz0[r] * (t - B); s2 += t - B;
          }
        } else {
          CT h = 0;
          const V2<CT> xc = x0[r];
          auto hf = [&](CT kk, CT th, V2<CT> xnb) {
            h += kk * ((th * xc.a + xc.b) - (th * xnb.a + xnb.b));
          };
          hf(kx[mW], tx[mW], xW);
          hf(kx[c], tx[c], xE);
          hf(kys, tys, xs);
          hf(kyc[r], tyc[r], xn);
          {
            const CT zc = tzm[r] * xc.a + xc.b, zn = tzm[r] * xm[r].a + xm[r].b;
            h += kzm[r] * (zc - zn);
            s2 += gzm[r] * v0[r] * (zc - zn);
          }
          {
            const CT zc = tz0[r] * xc.a + xc.b, zn = tz0[r] * xp[r].a + xp[r].b;
            h += kz0[r] * (zc - zn);
            s2 += gz0[r] * v0[r] * (zn - zc);
          }
          s1 += d0[r] * h;
        }
        double yv1, yv2;
        if constexpr ((FL & AF_B) != 0) {
          const bool dead = v0[r] == (CT)0 &&
                            kx[mW] + kx[c] + kys + kyc[r] + kzm[r] + kz0[r] == (CT)0;
          yv1 = dead ? 0.0 : b0.x + (double)s1;
          yv2 = dead ? 0.0 : b0.y + (double)s2;
        } else {
          yv1 = -(double)s1; yv2 = -(double)s2;
        }
        if (ok[r]) {
          if constexpr ((FL & AF_DX) != 0) st2(XW, c, (double)x0[r].a, (double)x0[r].b);
          st2(Y, c, yv1, yv2);
          if constexpr ((FL & AF_Y2) != 0) st2(Y2, c, yv1, yv2);
          if constexpr ((FL & AF_YY) != 0) acc[0] += yv1 * yv1 + yv2 * yv2;
          if constexpr ((FL & AF_YU) != 0) acc[1] += yv1 * (double)u0.a + yv2 * (double)u0.b;
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        xm[r] = x0[r]; dm[r] = d0[r]; vm[r] = v0[r]; kzm[r] = kz0[r]; tzm[r] = tz0[r];
        gzm[r] = gz0[r]; x0[r] = xp[r]; d0[r] = dp[r]; v0[r] = vp[r];
        xw0[r] = xwp[r]; xe0[r] = xep[r]; dw0[r] = dwp[r]; de0[r] = dep[r];
      }
      xS0 = xSp; dS0 = dSp; xN1 = xNp; dN1 = dNp;
    }
  }
}

// march_rows for 2x2 blocks with the loads two levels deep: iteration k issues the cell
// data (x, DUT, volume; west/east neighbours; halo rows) of level k + 2 and the face
// coefficients and dot operands of level k + 1, then computes level k from registers
// filled one iteration earlier.  Two iterations of loads are in flight per thread instead
// of one.  The arithmetic per cell is march_rows' (the same bits).  Levels past the top
// are read at the top level: in bounds, and the top z-face coefficients are zero.
template <int TR, int FL, int R, typename CT, typename VX, typename VO, typename VU>
__device__ __forceinline__ void march_rows_deep(const KApply<CT, VX, VO, VU>& a, int L,
                                                double (&acc)[2]) {
#pragma clang fp contract(on)  // as in march_rows
  static_assert((FL & AF_X0) == 0, "the x = 0 form keeps march_rows' walk");
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const long sy = nx, sz = (long)nx * ny;
  const CT* __restrict__ kx = a.c.kx;
  const CT* __restrict__ ky = a.c.ky;
  const CT* __restrict__ kz = a.c.kz;
  const CT* __restrict__ tx = a.c.tx;
  const CT* __restrict__ ty = a.c.ty;
  const CT* __restrict__ tz = a.c.tz;
  const CT* __restrict__ gzz = a.c.gz;
  const CT* __restrict__ vol = a.c.vol;
  const CT* __restrict__ dut = a.c.dut;
  const VX* __restrict__ X = a.x;
  const float* __restrict__ DX = a.dx;
  VO* __restrict__ Y = a.y;
  VO* __restrict__ Y2 = a.y2;
  const VU* __restrict__ U = a.u;
  double* __restrict__ XW = a.xw;
  const double* __restrict__ Bv = a.b;
  const CT fa = a.fa;
  if (L >= a.ntiles) return;
  const int bx = L % a.gx, rest = L / a.gx, by = rest % a.gy, kt = rest / a.gy;
  const int i = bx * a.bxw + threadIdx.x;
  const int j0 = (by * a.byw + threadIdx.y) * R;
  if (i >= nx || j0 >= ny) return;
  const int k0 = kt * a.kp, k1 = min(nz, k0 + a.kp);
  long cb[R];  // cell index of each row at level 0 (rows past ny: row ny-1)
  bool ok[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    ok[r] = j0 + r < ny;
    cb[r] = i + sy * (ok[r] ? j0 + r : ny - 1);
  }
  const long cbS = j0 > 0 ? cb[0] - sy : cb[0];
  const long cbN = j0 + R < ny ? cb[R - 1] + sy : cb[R - 1];
  const long dW = i > 0 ? -1 : 0, dE = i + 1 < nx ? 1 : 0;
  auto LX = [&](long m) -> V2<CT> {
    V2<VX> v = ld2(X, m);
    V2<CT> r{(CT)v.a, (CT)v.b};
    if constexpr ((FL & AF_DX) != 0) {
      V2<float> e = ld2(DX, m);
      r.a += (CT)e.a; r.b += (CT)e.b;
    }
    return r;
  };
  auto lev = [&](int k) -> long { return sz * (long)(k < nz ? k : nz - 1); };
  struct Cells {
    V2<CT> x[R], xw[R], xe[R], xS, xN;
    CT d[R], v[R], dw[R], de[R], dS, dN;
  };
  struct Faces {
    CT kz[R], tz[R], gz[R], ky[R], ty[R], kxw[R], kxe[R], txw[R], txe[R], kyS, tyS;
    V2<VU> u[R];
    double2 b[R];
  };
  auto ldcells = [&](int k, Cells& C) {
    const long o = lev(k);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const long c = cb[r] + o;
      C.x[r] = LX(c); C.d[r] = dut[c]; C.v[r] = ld_once(vol, c);
      C.xw[r] = LX(c + dW); C.xe[r] = LX(c + dE); C.dw[r] = 0; C.de[r] = 0;
      if constexpr (TR == 0) { C.dw[r] = dut[c + dW]; C.de[r] = dut[c + dE]; }
    }
    C.xS = LX(cbS + o); C.xN = LX(cbN + o); C.dS = 0; C.dN = 0;
    if constexpr (TR == 0) { C.dS = dut[cbS + o]; C.dN = dut[cbN + o]; }
  };
  auto ldfaces = [&](int k, Faces& F) {
    const long o = lev(k);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const long c = cb[r] + o;
      F.kz[r] = ld_once(kz, c); F.tz[r] = ld_once(tz, c); F.gz[r] = ld_once(gzz, c);
      F.ky[r] = ky[c]; F.ty[r] = ty[c];
      F.kxw[r] = kx[c + dW]; F.kxe[r] = kx[c]; F.txw[r] = tx[c + dW]; F.txe[r] = tx[c];
      F.u[r] = V2<VU>{0, 0};
      if constexpr ((FL & AF_YU) != 0) F.u[r] = ld2(U, c);
      F.b[r] = make_double2(0.0, 0.0);
      if constexpr ((FL & AF_B) != 0) F.b[r] = *reinterpret_cast<const double2*>(Bv + 2 * c);
    }
    F.kyS = ky[cbS + o]; F.tyS = ty[cbS + o];
  };
  V2<CT> xm[R];
  CT dm[R], vm[R], kzm[R], tzm[R], gzm[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    xm[r] = V2<CT>{0, 0}; dm[r] = 0; vm[r] = 0; kzm[r] = 0; tzm[r] = 0; gzm[r] = 0;
    if (k0 > 0) {
      const long m = cb[r] + sz * (k0 - 1);
      xm[r] = LX(m); dm[r] = dut[m]; vm[r] = ld_once(vol, m); kzm[r] = ld_once(kz, m);
      tzm[r] = ld_once(tz, m); gzm[r] = ld_once(gzz, m);
    }
  }
  Cells C0, C1, C2;
  Faces F0, F1;
  ldcells(k0, C0);
  ldcells(k0 + 1, C1);
  ldfaces(k0, F0);
  for (int k = k0; k < k1; ++k) {
    ldcells(k + 2, C2);
    ldfaces(k + 1, F1);
    const long off = sz * k;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const long c = cb[r] + off;
      const V2<CT> x0 = C0.x[r], xp = C1.x[r];
      const CT d0 = C0.d[r], v0 = C0.v[r], dp = C1.d[r], vp = C1.v[r];
      const CT kz0 = F0.kz[r], tz0 = F0.tz[r], gz0 = F0.gz[r];
      const V2<CT> xW = C0.xw[r], xE = C0.xe[r];
      const V2<CT> xs = r > 0 ? C0.x[r - 1] : C0.xS, xn = r + 1 < R ? C0.x[r + 1] : C0.xN;
      const CT ds = r > 0 ? C0.d[r - 1] : C0.dS, dn = r + 1 < R ? C0.d[r + 1] : C0.dN;
      const CT kys = r > 0 ? F0.ky[r - 1] : F0.kyS, tys = r > 0 ? F0.ty[r - 1] : F0.tyS;
      const V2<VU> u0 = F0.u[r];
      const double2 b0 = F0.b[r];
      CT s1 = fa * v0 * x0.a, s2 = v0 * x0.b;
      if constexpr (TR == 0) {
        const CT q0 = d0 * x0.a;
        auto hf = [&](CT kk, CT th, CT dnb, CT xnb) {
          const CT t = kk * (q0 - dnb * xnb);
          s1 += th * t; s2 += t;
        };
        hf(F0.kxw[r], F0.txw[r], C0.dw[r], xW.a);
        hf(F0.kxe[r], F0.txe[r], C0.de[r], xE.a);
        hf(kys, tys, ds, xs.a);
        hf(F0.ky[r], F0.ty[r], dn, xn.a);
        {
          const CT t = kzm[r] * (q0 - dm[r] * xm[r].a);
          const CT B = gzm[r] * (vm[r] * xm[r].b + v0 * x0.b);
          s1 += tzm[r] * (t + B); s2 += t + B;
        }
        {
          const CT t = kz0 * (q0 - dp * xp.a);
          const CT B = gz0 * (v0 * x0.b + vp * xp.b);
          s1 += tz0 * (t - B); s2 += t - B;
        }
      } else {
        CT h = 0;
        const V2<CT> xc = x0;
        auto hf = [&](CT kk, CT th, V2<CT> xnb) {
          h += kk * ((th * xc.a + xc.b) - (th * xnb.a + xnb.b));
        };
        hf(F0.kxw[r], F0.txw[r], xW);
        hf(F0.kxe[r], F0.txe[r], xE);
        hf(kys, tys, xs);
        hf(F0.ky[r], F0.ty[r], xn);
        {
          const CT zc = tzm[r] * xc.a + xc.b, zn = tzm[r] * xm[r].a + xm[r].b;
          h += kzm[r] * (zc - zn);
          s2 += gzm[r] * v0 * (zc - zn);
        }
        {
          const CT zc = tz0 * xc.a + xc.b, zn = tz0 * xp.a + xp.b;
          h += kz0 * (zc - zn);
          s2 += gz0 * v0 * (zn - zc);
        }
        s1 += d0 * h;
      }
      double yv1, yv2;
      if constexpr ((FL & AF_B) != 0) {
        const bool dead = v0 == (CT)0 &&
                          F0.kxw[r] + F0.kxe[r] + kys + F0.ky[r] + kzm[r] + kz0 == (CT)0;
        yv1 = dead ? 0.0 : b0.x + (double)s1;
        yv2 = dead ? 0.0 : b0.y + (double)s2;
      } else {
        yv1 = -(double)s1; yv2 = -(double)s2;
      }
      if (ok[r]) {
        if constexpr ((FL & AF_DX) != 0) st2(XW, c, (double)x0.a, (double)x0.b);
        st2(Y, c, yv1, yv2);
        if constexpr ((FL & AF_Y2) != 0) st2(Y2, c, yv1, yv2);
        if constexpr ((FL & AF_YY) != 0) acc[0] += yv1 * yv1 + yv2 * yv2;
        if constexpr ((FL & AF_YU) != 0) acc[1] += yv1 * (double)u0.a + yv2 * (double)u0.b;
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      xm[r] = C0.x[r]; dm[r] = C0.d[r]; vm[r] = C0.v[r];
      kzm[r] = F0.kz[r]; tzm[r] = F0.tz[r]; gzm[r] = F0.gz[r];
    }
    C0 = C1;
    C1 = C2;
    F0 = F1;
  }
}

}  // namespace asam
