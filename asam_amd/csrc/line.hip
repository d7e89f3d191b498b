// line.hip -- vertical-line preconditioner solve (gfx950): one thread per column, 64
// columns per wavefront (coalesced along x at every level), loads issued a window of
// levels ahead of the Thomas recurrence, forward-sweep results staged in LDS.
#include "line_tile.h"

namespace asam {

// ===================================================================== line solve
int line_blocks(const Geom& g) { return ((g.nx + 63) / 64) * g.ny; }

// Dynamic LDS a line-solve workgroup may take: the device's per-workgroup limit minus
// 4 KiB kept for the static reduction scratch.
size_t lds_device_limit() {
  static int cached = -1;
  if (cached < 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      v = 0;
    cached = v > 4096 ? v - 4096 : 0;
  }
  return (size_t)cached;
}

// Default staging budget per workgroup: at most 60 KiB, so that two or more one-wavefront
// workgroups share a CU.
size_t lds_budget() {
  const size_t d = lds_device_limit();
  return d < 60 * 1024 ? d : 60 * 1024;
}

// Staging budget the line_lds tuning knob names: 0 none, 1 the default, n >= 2 n KiB
// (never more than the device allows).
size_t lds_budget_for(int line_lds) {
  if (line_lds <= 0) return 0;
  if (line_lds == 1) return lds_budget();
  const size_t want = (size_t)line_lds * 1024, d = lds_device_limit();
  return want < d ? want : d;
}

// Levels of a column of nz levels staged in LDS within `budget` bytes.
int lds_levels(int nz, int np, size_t word, size_t budget) {
  const size_t per = (size_t)64 * np * word;
  const size_t k = budget / per;
  return k >= (size_t)nz ? nz : (int)k;
}

// Backward-sweep window: 16 levels with the 8-level forward window (the backward sweep
// holds 6-10 words per level against the forward's 12-14; measured 1-5% faster line
// solves), else the forward window.
template <int P>
constexpr int kBackWindow = P == 8 ? 16 : 0;

// One 64-column tile per workgroup: grid (gxl, ny).
template <int NP, int TR, int MODE, int LDS, int P, int NT, typename CT, typename VT>
__global__ __launch_bounds__(64) void k_line(KLine<CT, VT> a) {
  if (a.done && *a.done) return;  // the state before this launch: the same in every workgroup
  extern __shared__ unsigned char lds_raw[];
  CT* stage = reinterpret_cast<CT*>(lds_raw);
  const int blk = blockIdx.x + gridDim.x * blockIdx.y;
  double acc[1] = {0.0};
  if constexpr (MODE == LINE_PLAIN) {
    line_tile<NP, TR, MODE, LDS, P, CT, VT, NT, kBackWindow<P>>(a, blockIdx.x, blockIdx.y,
                                                                gridDim.x, stage, acc);
  } else {
    // a folded tail (the previous iteration's exit test and beta before the p-step, alpha
    // before the s-step): the scalars come from this workgroup's copy of the state
    __shared__ unsigned long long sk[kTailWords];
    const bool fold = a.fold.kind >= 0;
    bool run = true;
    KLine<CT, VT> b = a;
    if (fold) {
      fold_state_w64(a.fold, sk);
      fold_write(a.fold, sk, threadIdx.x, 64);
      const KState* s = reinterpret_cast<const KState*>(sk);
      b.st = s;
      run = s->done == 0;
    }
    if (run)
      line_tile<NP, TR, MODE, LDS, P, CT, VT, NT, kBackWindow<P>>(b, blockIdx.x, blockIdx.y,
                                                                  gridDim.x, stage, acc);
    if constexpr (MODE == LINE_BICG_S) {
      block_reduce_store<1>(acc, a.partial + (long)a.qslot * kRedBlocks, blk, threadIdx.x, 64);
    }
  }
}

template <int NP, int TR, int LDS, int P, int NT, typename CT, typename VT>
static void line_mode(hipStream_t s, const Geom& g, const KLine<CT, VT>& a, int mode,
                      size_t lds) {
  dim3 grid((g.nx + 63) / 64, g.ny);
  switch (mode) {
    case LINE_PLAIN:
      hipLaunchKernelGGL((k_line<NP, TR, LINE_PLAIN, LDS, P, NT, CT, VT>), grid, dim3(64), lds, s,
                         a);
      break;
    case LINE_BICG_P:
      hipLaunchKernelGGL((k_line<NP, TR, LINE_BICG_P, LDS, P, NT, CT, VT>), grid, dim3(64), lds,
                         s, a);
      break;
    default:
      hipLaunchKernelGGL((k_line<NP, TR, LINE_BICG_S, LDS, P, NT, CT, VT>), grid, dim3(64), lds,
                         s, a);
      break;
  }
}

// Cache policy of the factor loads: the tuning's, or by the factors' footprint.
template <int NP, typename CT>
static bool line_nontemporal(const Geom& g) {
  if (g.tune.line_nt != 2) return g.tune.line_nt == 1;
  return 3l * NP * NP * fac_cells(g) * (long)sizeof(CT) > kLineNtBytes;
}

template <int NP, int TR, int LDS, int P, typename CT, typename VT>
static void line_launch(hipStream_t s, const Geom& g, const KLine<CT, VT>& a, int mode,
                        size_t lds) {
  if (line_nontemporal<NP, CT>(g)) line_mode<NP, TR, LDS, P, 1, CT, VT>(s, g, a, mode, lds);
  else line_mode<NP, TR, LDS, P, 0, CT, VT>(s, g, a, mode, lds);
}

template <int NP, int TR, typename CT, typename VT>
static void line_t(hipStream_t s, const Geom& g, const Coef<CT>& c, const LineReq& r) {
  KLine<CT, VT> a;
  a.c = c; a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.tiled = g.tune.line_tiled;
  a.src = (const VT*)r.src; a.p = (VT*)r.p; a.v = (const VT*)r.v; a.out = (VT*)r.out;
  a.scale = r.scale; a.st = r.st; a.partial = r.partial; a.done = r.done;
  a.qslot = r.qslot;
  a.fold = r.fold;
  a.kl = r.use_lds ? lds_levels(g.nz, NP, sizeof(CT), lds_budget_for(g.tune.line_lds)) : 0;
  const size_t lds = (size_t)a.kl * 64 * NP * sizeof(CT);
  const bool use_lds = a.kl > 0;
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
