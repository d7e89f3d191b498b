// apply.hip -- stencil apply kernels (gfx950).
//
// The operator of one tile is march_tile (apply_tile.h).
//
// Kernel shapes: apply = one cell per thread over 64x4-column tiles of kp planes
// (neighbour re-reads hit L1/L2; the tiles of an XCD are contiguous so its L2 holds
// the planes above and below).  Line solve = one thread per vertical column, 64
// columns per wavefront (coalesced along x at every level), loads issued a window of
// levels ahead of the recurrence, forward-sweep results staged in LDS for the
// backward sweep.
#include "asam_sound.h"
#include "apply_tile.h"

namespace asam {

struct ApplyShape {
  int gx, gy, kp, ntiles, nblocks;
};

// Rows per band of the band-ordered variant: the tuning's value, or ny/8 rounded up so that
// each XCD owns one band.
static int band_rows(const Geom& g) {
  const int byw = 256 / g.tune.apply_bx;
  int rb = g.tune.apply_band_rows > 0 ? g.tune.apply_band_rows : (g.ny + 7) / 8;
  if (rb < byw) rb = byw;
  return rb;
}

// Variants 0, 1 and 3: tiles of bx x (256/bx) columns (variant 3: x apply_rows rows per
// thread) by kp planes, at most ~apply_wg
// workgroups, padded to a multiple of the 8 XCDs so that the physical -> logical remap below
// is a bijection.  Variant 2: apply_wg workgroups (a multiple of 8) looping over the tiles
// of their XCD's bands.
static ApplyShape apply_shape(const Geom& g, int fp32) {
  ApplyShape a;
  const int bxw = g.tune.apply_bx, byw = 256 / bxw;
  a.gx = (g.nx + bxw - 1) / bxw;
  if (g.tune.apply_variant == 2) {
    const int rb = band_rows(g);
    a.gy = (rb + byw - 1) / byw;  // row tiles per band
    a.kp = 1;
    a.ntiles = 0;
    int nb = (g.tune.apply_wg / 8) * 8;
    if (nb < 8) nb = 8;
    a.nblocks = nb;
    return a;
  }
  // variant 3 tiles: four rows per thread with FP32 coefficients only (see apply_t)
  const int rpt = (g.tune.apply_variant == 3 && g.tune.apply_rows == 4 && fp32) ? 4 : 2;
  const int rows = g.tune.apply_variant == 3 ? byw * rpt : byw;
  a.gy = (g.ny + rows - 1) / rows;
  long cols = (long)a.gx * a.gy;
  long want = g.tune.apply_wg;
  int kp = (int)((cols * g.nz + want - 1) / want);
  if (kp < 1) kp = 1;
  if (kp > g.nz) kp = g.nz;
  a.kp = kp;
  int nzt = (g.nz + kp - 1) / kp;
  a.ntiles = (int)(cols * nzt);
  a.nblocks = ((a.ntiles + 7) / 8) * 8;
  return a;
}

int apply_blocks(const Geom& g, int fp32) { return apply_shape(g, fp32).nblocks; }

// The operator at one cell (i, j, k) with the fused options (k_apply and k_apply_band).
// Neighbour indices are clamped to the cell itself at walls: the face difference
// (q_c - q_n) is then exactly zero, so no branch separates the loads.
template <int NP, int TR, int FL, typename CT, typename VX, typename VO, typename VU>
__device__ __forceinline__ void apply_cell(const KApply<CT, VX, VO, VU>& a, int i, int j, int k,
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
  const CT fa = a.fa;
  const long c = i + sy * j + sz * k;
  if constexpr ((FL & AF_X0) != 0) {
    // x = 0: every stencil term is a signed zero, so b + s equals b; same accumulation
    // order as the full form
    static_assert((FL & AF_B) != 0 && (FL & (AF_DX | AF_YU)) == 0, "x = 0 residual form");
    const bool dead = a.dead[c] != 0;
    if constexpr (NP == 1) {
      const double yv = dead ? 0.0 : a.b[c];
      a.y[c] = (VO)yv;
      if constexpr ((FL & AF_Y2) != 0) a.y2[c] = (VO)yv;
      if constexpr ((FL & AF_YY) != 0) acc[0] += yv * yv;
    } else {
      const double2 b0 = *reinterpret_cast<const double2*>(a.b + 2 * c);
      const double yv1 = dead ? 0.0 : b0.x, yv2 = dead ? 0.0 : b0.y;
      st2(a.y, c, yv1, yv2);
      if constexpr ((FL & AF_Y2) != 0) st2(a.y2, c, yv1, yv2);
      if constexpr ((FL & AF_YY) != 0) acc[0] += yv1 * yv1 + yv2 * yv2;
    }
    return;
  }
  const long mW = i > 0 ? c - 1 : c, mE = i + 1 < nx ? c + 1 : c;
  const long mS = j > 0 ? c - sy : c, mN = j + 1 < ny ? c + sy : c;
  const long mD = k > 0 ? c - sz : c, mU = k + 1 < nz ? c + sz : c;
  double yv1, yv2 = 0.0;
  if constexpr (NP == 1) {
    auto LX = [&](long m) -> CT {
      CT v = (CT)X[m];
      if constexpr ((FL & AF_DX) != 0) v += (CT)DX[m];
      return v;
    };
    const CT x0 = LX(c), d0 = dut[c];
    CT s = fa * vol[c] * x0;
    if constexpr (TR == 0) {
      const CT q0 = d0 * x0;
      s += kx[mW] * (q0 - dut[mW] * LX(mW));
      s += kx[c] * (q0 - dut[mE] * LX(mE));
      s += ky[mS] * (q0 - dut[mS] * LX(mS));
      s += ky[c] * (q0 - dut[mN] * LX(mN));
      s += kz[mD] * (q0 - dut[mD] * LX(mD));
      s += kz[c] * (q0 - dut[mU] * LX(mU));
    } else {
      CT h = kx[mW] * (x0 - LX(mW));
      h += kx[c] * (x0 - LX(mE));
      h += ky[mS] * (x0 - LX(mS));
      h += ky[c] * (x0 - LX(mN));
      h += kz[mD] * (x0 - LX(mD));
      h += kz[c] * (x0 - LX(mU));
      s += d0 * h;
    }
    // Op = -A x; residual form y = b - Op(x) = b + A x, zero on dead cells
    if constexpr ((FL & AF_B) != 0) {
      const bool dead = vol[c] == (CT)0 && kx[mW] + kx[c] + ky[mS] + ky[c] +
                                           kz[mD] + kz[c] == (CT)0;
      yv1 = dead ? 0.0 : a.b[c] + (double)s;
    } else {
      yv1 = -(double)s;
    }
    if constexpr ((FL & AF_DX) != 0) a.xw[c] = (double)x0;
    a.y[c] = (VO)yv1;
    if constexpr ((FL & AF_Y2) != 0) a.y2[c] = (VO)yv1;
    if constexpr ((FL & AF_YY) != 0) acc[0] += yv1 * yv1;
    if constexpr ((FL & AF_YU) != 0) acc[1] += yv1 * (double)ld_once(a.u, c);
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
    const V2<CT> x0 = LX(c);
    const CT d0 = dut[c], v0 = vol[c];
    const V2<CT> xW = LX(mW), xE = LX(mE), xS = LX(mS), xN = LX(mN);
    const V2<CT> xD = LX(mD), xU = LX(mU);
    // gravity couplings of the z-faces below and above (zero at the walls)
    const CT gD = k > 0 ? gzz[mD] : (CT)0, gU = gzz[c];
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
        const CT t = kz[mD] * (q0 - dut[mD] * xD.a);
        const CT B = gD * (vol[mD] * xD.b + v0 * x0.b);
        s1 += tz[mD] * (t + B); s2 += t + B;
      }
      {
        const CT t = kz[c] * (q0 - dut[mU] * xU.a);
        const CT B = gU * (v0 * x0.b + vol[mU] * xU.b);
        s1 += tz[c] * (t - B); s2 += t - B;
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
        const CT th = tz[mD];
        const CT zc = th * x0.a + x0.b, zn = th * xD.a + xD.b;
        h += kz[mD] * (zc - zn);
        s2 += gD * v0 * (zc - zn);
      }
      {
        const CT th = tz[c];
        const CT zc = th * x0.a + x0.b, zn = th * xU.a + xU.b;
        h += kz[c] * (zc - zn);
        s2 += gU * v0 * (zn - zc);
      }
      s1 += d0 * h;
    }
    if constexpr ((FL & AF_B) != 0) {
      const double2 bb = *reinterpret_cast<const double2*>(a.b + 2 * c);
      const bool dead = v0 == (CT)0 && kx[mW] + kx[c] + ky[mS] + ky[c] +
                                           kz[mD] + kz[c] == (CT)0;
      yv1 = dead ? 0.0 : bb.x + (double)s1;
      yv2 = dead ? 0.0 : bb.y + (double)s2;
    } else {
      yv1 = -(double)s1; yv2 = -(double)s2;
    }
    if constexpr ((FL & AF_DX) != 0) st2(a.xw, c, (double)x0.a, (double)x0.b);
    st2(a.y, c, yv1, yv2);
    if constexpr ((FL & AF_Y2) != 0) st2(a.y2, c, yv1, yv2);
    if constexpr ((FL & AF_YY) != 0) acc[0] += yv1 * yv1 + yv2 * yv2;
    if constexpr ((FL & AF_YU) != 0) {
      const V2<VU> uu = ld2(a.u, c);
      acc[1] += yv1 * (double)uu.a + yv2 * (double)uu.b;
    }
  }
}

template <int NP, int TR, int FL, typename CT, typename VX, typename VO, typename VU>
__global__ __launch_bounds__(256) void k_apply(KApply<CT, VX, VO, VU> a) {
  if (a.done && *a.done) return;
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a contiguous
  // range of tiles (consecutive z-planes of the same columns) so that z- and
  // y-neighbour re-reads hit that XCD's L2.
  const int bphys = blockIdx.x, nb8 = gridDim.x >> 3;
  const int L = a.xcd ? (bphys & 7) * nb8 + (bphys >> 3) : bphys;
  double acc[2] = {0.0, 0.0};
  if (L < a.ntiles) {
    const int bx = L % a.gx, rest = L / a.gx, by = rest % a.gy, kt = rest / a.gy;
    const int i = bx * a.bxw + threadIdx.x, j = by * a.byw + threadIdx.y;
    const int k0 = kt * a.kp, k1 = min(a.nz, k0 + a.kp);
    if (i < a.nx && j < a.ny) {
#pragma unroll 2
      for (int k = k0; k < k1; ++k) apply_cell<NP, TR, FL, CT, VX, VO, VU>(a, i, j, k, acc);
    }
  }
  if constexpr ((FL & (AF_YY | AF_YU)) != 0) {
    block_reduce_store<2>(acc, a.partial, bphys, tid, 256);
    if (a.tail.kind >= 0) tail_arrive(a.tail, tid);
  }
}

// Band-ordered flat variant: one cell per thread.  XCD x owns the bands x, x+8, x+16, ...
// of band_rows rows; its workgroups walk one band at a time, plane by plane, tile after tile
// (bx x by cells of one plane, x fastest), each workgroup taking every G-th tile of the band.
// The G workgroups of an XCD so work on a window of consecutive tiles: the rows above and
// below and the plane below were fetched by the same XCD moments earlier and hit its L2, and
// HBM sees every array streamed once in address order per band.
template <int NP, int TR, int FL, typename CT, typename VX, typename VO, typename VU>
__global__ __launch_bounds__(256) void k_apply_band(KApply<CT, VX, VO, VU> a) {
  if (a.done && *a.done) return;
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int bphys = blockIdx.x, xcd = bphys & 7, slot = bphys >> 3, G = gridDim.x >> 3;
  const int rb = a.band_rows, nbands = (a.ny + rb - 1) / rb;
  const int tpp = a.gx * a.gy;            // tiles per plane of a band
  const long tpb = (long)tpp * a.nz;      // tiles per band
  double acc[2] = {0.0, 0.0};
  for (int band = xcd; band < nbands; band += 8) {
    const int j0 = band * rb, j1 = min(a.ny, j0 + rb);
    for (long t = slot; t < tpb; t += G) {
      const int k = (int)(t / tpp), r = (int)(t - (long)k * tpp);
      const int ty = r / a.gx, tx = r - ty * a.gx;
      const int i = tx * a.bxw + threadIdx.x, j = j0 + ty * a.byw + threadIdx.y;
      if (i < a.nx && j < j1) apply_cell<NP, TR, FL, CT, VX, VO, VU>(a, i, j, k, acc);
    }
  }
  if constexpr ((FL & (AF_YY | AF_YU)) != 0) {
    block_reduce_store<2>(acc, a.partial, bphys, tid, 256);
    if (a.tail.kind >= 0) tail_arrive(a.tail, tid);
  }
}

// z-march variant: one tile per workgroup (tiles of an XCD contiguous, see k_apply).
template <int NP, int TR, int FL, typename CT, typename VX, typename VO, typename VU>
__global__ __launch_bounds__(256) void k_apply_march(KApply<CT, VX, VO, VU> a) {
  if (a.done && *a.done) return;
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int bphys = blockIdx.x, nb8 = gridDim.x >> 3;
  const int L = a.xcd ? (bphys & 7) * nb8 + (bphys >> 3) : bphys;
  double acc[2] = {0.0, 0.0};
  march_tile<NP, TR, FL, CT, VX, VO, VU>(a, L, acc);
  if constexpr ((FL & (AF_YY | AF_YU)) != 0) {
    block_reduce_store<2>(acc, a.partial, bphys, tid, 256);
    if (a.tail.kind >= 0) tail_arrive(a.tail, tid);
  }
}

// Row-blocked z-march variant: R rows per thread (march_rows; DEEP 1: march_rows_deep, loads
// two levels ahead; DEEP 2: march_rows reading the packed z-face coefficients and volume),
// tiles of an XCD contiguous.
template <int NP, int TR, int FL, int R, typename CT, typename VX, typename VO, typename VU,
          int DEEP = 0>
__global__ __launch_bounds__(256) void k_apply_rows(KApply<CT, VX, VO, VU> a) {
  if (a.done && *a.done) return;
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int bphys = blockIdx.x, nb8 = gridDim.x >> 3;
  const int L = a.xcd ? (bphys & 7) * nb8 + (bphys >> 3) : bphys;
  double acc[2] = {0.0, 0.0};
  if constexpr (DEEP == 1) march_rows_deep<TR, FL, R, CT, VX, VO, VU>(a, L, acc);
  else if constexpr (DEEP == 2) march_rows<NP, TR, FL, R, CT, VX, VO, VU, 1>(a, L, acc);
  else march_rows<NP, TR, FL, R, CT, VX, VO, VU>(a, L, acc);
  if constexpr ((FL & (AF_YY | AF_YU)) != 0) {
    block_reduce_store<2>(acc, a.partial, bphys, tid, 256);
    if (a.tail.kind >= 0) tail_arrive(a.tail, tid);
  }
}

template <int NP, int TR, int FL, typename CT, typename VX, typename VO, typename VU>
static void apply_t(hipStream_t s, const Geom& g, const Coef<CT>& c, const ApplyReq& r) {
  KApply<CT, VX, VO, VU> a;
  a.c = c; a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.fa = (CT)g.fa;
  ApplyShape sh = apply_shape(g, sizeof(CT) == 4 ? 1 : 0);
  a.gx = sh.gx; a.gy = sh.gy; a.kp = sh.kp; a.ntiles = sh.ntiles;
  a.bxw = g.tune.apply_bx; a.byw = 256 / g.tune.apply_bx; a.xcd = g.tune.apply_xcd;
  a.x = (const VX*)r.x; a.dx = r.dx; a.xw = r.xw; a.b = r.b; a.y = (VO*)r.y;
  a.y2 = (VO*)r.y2; a.u = (const VU*)r.u; a.partial = r.partial; a.done = r.done;
  a.dead = r.x0_dead;
  a.tail = r.tail;
  a.band_rows = band_rows(g);
  if (g.tune.apply_variant == 2) {
    // the x = 0 form too, so that it writes the partial sums of the full form
    hipLaunchKernelGGL((k_apply_band<NP, TR, FL, CT, VX, VO, VU>), dim3(sh.nblocks),
                       dim3(a.bxw, a.byw), 0, s, a);
  } else if (g.tune.apply_variant == 3) {
    // the x = 0 form too (same walk, so the same partial sums as the full form).  Four rows
    // only with FP32 coefficients: FP64 kernels of four rows spill registers, and the FP64
    // applies of a mixed solve share the FP32 inner solve's tuning.
    bool four = false;
    if constexpr (sizeof(CT) == 4) four = g.tune.apply_rows == 4;
    // loads two levels ahead: FP32 coefficients, 2x2 blocks and two rows (the FP32 inner
    // solve's applies); FP64 kernels and four rows would spill the extra registers
    if constexpr (sizeof(CT) == 4 && NP == 2 && (FL & AF_X0) == 0) {
      if (g.tune.apply_depth == 2 && !four) {
        hipLaunchKernelGGL((k_apply_rows<NP, TR, FL, 2, CT, VX, VO, VU, 1>), dim3(sh.nblocks),
                           dim3(a.bxw, a.byw), 0, s, a);
        return;
      }
      // one level ahead: the z-face coefficients and volume from their packed copy
      if (c.pkz) {
        if (four)
          hipLaunchKernelGGL((k_apply_rows<NP, TR, FL, 4, CT, VX, VO, VU, 2>), dim3(sh.nblocks),
                             dim3(a.bxw, a.byw), 0, s, a);
        else
          hipLaunchKernelGGL((k_apply_rows<NP, TR, FL, 2, CT, VX, VO, VU, 2>), dim3(sh.nblocks),
                             dim3(a.bxw, a.byw), 0, s, a);
        return;
      }
    }
    if (four) {
      if constexpr (sizeof(CT) == 4)
        hipLaunchKernelGGL((k_apply_rows<NP, TR, FL, 4, CT, VX, VO, VU>), dim3(sh.nblocks),
                           dim3(a.bxw, a.byw), 0, s, a);
    } else {
      hipLaunchKernelGGL((k_apply_rows<NP, TR, FL, 2, CT, VX, VO, VU>), dim3(sh.nblocks),
                         dim3(a.bxw, a.byw), 0, s, a);
    }
  } else if constexpr ((FL & AF_X0) != 0) {
    // same grid and partial-sum slots in either variant
    hipLaunchKernelGGL((k_apply_march<NP, TR, FL, CT, VX, VO, VU>), dim3(sh.nblocks),
                       dim3(a.bxw, a.byw), 0, s, a);
  } else if (g.tune.apply_variant == 0) {
    hipLaunchKernelGGL((k_apply_march<NP, TR, FL, CT, VX, VO, VU>), dim3(sh.nblocks),
                       dim3(a.bxw, a.byw), 0, s, a);
  } else {
    hipLaunchKernelGGL((k_apply<NP, TR, FL, CT, VX, VO, VU>), dim3(sh.nblocks), dim3(a.bxw, a.byw),
                       0, s, a);
  }
}

// The fused-option combinations the solvers use; any other set (or an x = 0 form with FP32
// coefficients) launches nothing and returns ASAM_E_INVALID.
template <int NP, int TR, typename CT, typename VX, typename VO, typename VU>
static int apply_fl(hipStream_t s, const Geom& g, const Coef<CT>& c, const ApplyReq& r) {
  int fl = (r.b ? AF_B : 0) | (r.y2 ? AF_Y2 : 0) | (r.dx ? AF_DX : 0) |
           ((r.dots & DOT_YY) ? AF_YY : 0) | ((r.dots & DOT_YU) ? AF_YU : 0) |
           (r.x0_dead ? AF_X0 : 0);
  switch (fl) {
    case AF_X0 | AF_B | AF_Y2 | AF_YY:
      if constexpr (sizeof(CT) == 8) {
        apply_t<NP, TR, AF_X0 | AF_B | AF_Y2 | AF_YY, CT, VX, VO, VU>(s, g, c, r);
        return 0;
      }
      return ASAM_E_INVALID;
    case AF_X0 | AF_B | AF_YY:
      if constexpr (sizeof(CT) == 8) {
        apply_t<NP, TR, AF_X0 | AF_B | AF_YY, CT, VX, VO, VU>(s, g, c, r);
        return 0;
      }
      return ASAM_E_INVALID;
    case AF_X0 | AF_B:
      if constexpr (sizeof(CT) == 8) {
        apply_t<NP, TR, AF_X0 | AF_B, CT, VX, VO, VU>(s, g, c, r);
        return 0;
      }
      return ASAM_E_INVALID;
    case 0: apply_t<NP, TR, 0, CT, VX, VO, VU>(s, g, c, r); return 0;
    case AF_YU: apply_t<NP, TR, AF_YU, CT, VX, VO, VU>(s, g, c, r); return 0;
    case AF_YY | AF_YU: apply_t<NP, TR, AF_YY | AF_YU, CT, VX, VO, VU>(s, g, c, r); return 0;
    case AF_B: apply_t<NP, TR, AF_B, CT, VX, VO, VU>(s, g, c, r); return 0;
    case AF_B | AF_YY: apply_t<NP, TR, AF_B | AF_YY, CT, VX, VO, VU>(s, g, c, r); return 0;
    case AF_B | AF_Y2 | AF_YY:
      apply_t<NP, TR, AF_B | AF_Y2 | AF_YY, CT, VX, VO, VU>(s, g, c, r);
      return 0;
    case AF_B | AF_Y2 | AF_DX | AF_YY:
      apply_t<NP, TR, AF_B | AF_Y2 | AF_DX | AF_YY, CT, VX, VO, VU>(s, g, c, r);
      return 0;
    default:
      return ASAM_E_INVALID;
  }
}

template <int NP, int TR>
static int apply_np(hipStream_t s, const Geom& g, const Coef<double>& c64,
                    const Coef<float>& c32, const ApplyReq& r) {
  if (r.prec == 0) {
    if (r.vo == 0) return apply_fl<NP, TR, double, double, double, double>(s, g, c64, r);
    return apply_fl<NP, TR, double, double, float, double>(s, g, c64, r);
  }
  return apply_fl<NP, TR, float, float, float, float>(s, g, c32, r);
}

int launch_apply(hipStream_t s, const Geom& g, const Coef<double>& c64, const Coef<float>& c32,
                 const ApplyReq& r) {
  if (g.np1 == 1) return r.tr ? apply_np<1, 1>(s, g, c64, c32, r) : apply_np<1, 0>(s, g, c64, c32, r);
  return r.tr ? apply_np<2, 1>(s, g, c64, c32, r) : apply_np<2, 0>(s, g, c64, c32, r);
}

}  // namespace asam
