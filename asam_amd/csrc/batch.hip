// batch.hip -- ensembles of uneven stage systems solved together (gfx950).
//
// Layout in HBM: the members' arrays are concatenated into one pool per array; member m's
// cells start at off[m] (coefficients), its column-tiled line factors at foff[m], its
// vector elements at np1*off[m].  Every kernel walks a table of tiles that spans all
// members (tile g belongs to member tmem[g]); the tiles of a member whose solve has
// finished are skipped one by one, so the work of a launch shrinks as members converge.
// The grid is either one workgroup per tile or persistent (pf workgroups per CU looping
// over the tiles); in both, workgroup b runs on XCD b & 7 (round-robin dispatch) and takes
// tiles from that XCD's contiguous eighth of the table, so neighbouring tiles of a member
// share an L2.
//
// Solve: FP32-inner BiCGStab with FP64 refinement per member (the single-domain mixed
// solve, same kernels' tile bodies and the same scalar tails), all members in lock step.
// The schedule (refinement steps x inner iterations) of one solve is captured into a HIP
// graph; kernels of members that are done return at once, so the schedule only has to be
// long enough.  A member that needs more than the schedule gives keeps iterating after
// the graph in eagerly enqueued steps (an inner solve cut short by the schedule is a valid
// refinement step: the FP64 outer residual is exact).
#include <limits.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "asam_ensemble.h"
#include "apply_tile.h"
#include "line_tile.h"

namespace asam {

size_t lds_budget();
size_t lds_budget_for(int line_lds);
int lds_levels(int nz, int np, size_t word, size_t budget);

struct EnsMember {
  int nx, ny, nz, np1;
  long off, n;                  // first cell in the pools, cells
  long foff;                    // first line-factor entry (column-tiled layout)
  int agx, agy, akp, ant, at0;  // apply tiles: agx x agy column tiles x (nz / akp), first
  int lgx, lnt, lt0;            // line tiles: lgx 64-column tiles per row x ny rows
  int bnt, bt0, pad;            // vector tiles of `chunk` elements
  double fa;
};

enum { SRC_A = 0, SRC_L = 1, SRC_B = 2 };  // partial-sum producers: apply, line, update

// -------------------------------------------------------------------- kernels
template <typename CT, typename VX, typename VO, typename VU>
struct KEnsApply {
  Coef<CT> c;
  const EnsMember* mem;
  const int* tmem;
  int ntiles, bxw;
  VX* xa;           // x (refinement: buffer 0 of the pair)
  VX* xb;           // refinement: buffer 1 of the pair
  const float* dx;
  const double* b;
  VO* y;
  VO* y2;
  const VU* u;
  double* partial;
  long pstride;
  const KState* st;
  int skip_outer;   // skip members with outer_done (else members with done)
  const unsigned char* dead = nullptr;  // AF_X0: cells whose residual is 0 (all members)
  const int* skip = nullptr;            // the whole launch returns at once when *skip != 0
};

// PAR: refinement apply x_eff = X[(o-1)&1] + d written to X[o&1], o = refinement step.
// RB: rows per thread (1: march_tile, 2: the row-blocked march_rows).
template <int NP, int FL, int PAR, int RB, typename CT, typename VX, typename VO, typename VU>
__global__ __launch_bounds__(256) void k_ens_apply(KEnsApply<CT, VX, VO, VU> e) {
  if (e.skip && *e.skip) return;
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int nb8 = gridDim.x >> 3, per = (e.ntiles + 7) >> 3;
  const int g0 = (blockIdx.x & 7) * per, g1 = min(e.ntiles, g0 + per);
  for (int g = g0 + (blockIdx.x >> 3); g < g1; g += nb8) {
    const int m = e.tmem[g];
    const KState* st = e.st + m;
    if (e.skip_outer ? st->outer_done : st->done) continue;
    const EnsMember M = e.mem[m];
    KApply<CT, VX, VO, VU> a;
    a.c = e.c;
    a.c.kx += M.off; a.c.ky += M.off; a.c.kz += M.off;
    a.c.vol += M.off; a.c.dut += M.off;
    if constexpr (NP == 2) { a.c.tx += M.off; a.c.ty += M.off; a.c.tz += M.off; a.c.gz += M.off; }
    if (a.c.pkz) a.c.pkz += 4 * M.off;
    a.nx = M.nx; a.ny = M.ny; a.nz = M.nz;
    a.gx = M.agx; a.gy = M.agy; a.kp = M.akp; a.ntiles = M.ant;
    a.bxw = e.bxw; a.byw = 256 / e.bxw; a.xcd = 0;
    a.fa = (CT)M.fa;
    const long v = (long)NP * M.off;
    VX* xs = e.xa;
    if constexpr (PAR) {
      const int o = st->outer;
      xs = ((o - 1) & 1) ? e.xb : e.xa;
      a.xw = (double*)(((o & 1) ? e.xb : e.xa) + v);
      a.dx = e.dx + v;
    } else {
      a.xw = nullptr;
      a.dx = nullptr;
    }
    a.x = xs + v;
    a.b = e.b ? e.b + v : nullptr;
    a.y = e.y + v;
    a.y2 = e.y2 ? e.y2 + v : nullptr;
    a.u = e.u ? e.u + v : nullptr;
    a.partial = nullptr;
    a.done = nullptr;
    a.dead = e.dead ? e.dead + M.off : nullptr;
    double acc[2] = {0.0, 0.0};
    // FP32 2x2 blocks: z-face coefficients and volume from the packed pool
    constexpr int ZP = NP == 2 && sizeof(CT) == 4 && (FL & AF_X0) == 0 ? 1 : 0;
    if constexpr (RB == 2) march_rows<NP, 0, FL, 2, CT, VX, VO, VU, ZP>(a, g - M.at0, acc);
    else march_tile<NP, 0, FL, CT, VX, VO, VU>(a, g - M.at0, acc);
    if constexpr ((FL & (AF_YY | AF_YU)) != 0) {
      block_reduce_store<2>(acc, e.partial, g, tid, 256, e.pstride);
      __syncthreads();  // the reduction scratch is reused by the next tile
    }
  }
}

template <typename CT, typename VT>
struct KEnsLine {
  Coef<CT> c;
  const EnsMember* mem;
  const int* tmem;
  int ntiles;
  const VT* src;
  VT* p;
  const VT* v;
  VT* out;
  const KState* st;
  double* partial;
  long pstride;
  int qslot;
  int kl;  // levels staged in LDS (the rest through the output vector)
};

template <int NP, int MODE, int LDS, int P, typename CT, typename VT>
__global__ __launch_bounds__(64) void k_ens_line(KEnsLine<CT, VT> e) {
  extern __shared__ unsigned char lds_raw[];
  CT* stage = reinterpret_cast<CT*>(lds_raw);
  const int nb8 = gridDim.x >> 3, per = (e.ntiles + 7) >> 3;
  const int g0 = (blockIdx.x & 7) * per, g1 = min(e.ntiles, g0 + per);
  for (int g = g0 + (blockIdx.x >> 3); g < g1; g += nb8) {
    const int m = e.tmem[g];
    const KState* st = e.st + m;
    if (st->done) continue;
    const EnsMember M = e.mem[m];
    KLine<CT, VT> a;
    a.c = e.c;
    const long fo = (long)NP * NP * M.foff;
    a.c.pl += lower_words(NP) * M.foff; a.c.pt += fo; a.c.pw += fo;
    a.nx = M.nx; a.ny = M.ny; a.nz = M.nz; a.tiled = 1;
    const long v = (long)NP * M.off;
    a.src = e.src ? e.src + v : nullptr;
    a.p = e.p + v;
    a.v = e.v ? e.v + v : nullptr;
    a.out = e.out + v;
    a.scale = nullptr; a.st = st; a.partial = nullptr; a.done = nullptr; a.qslot = 0;
    a.kl = e.kl;
    const int lt = g - M.lt0;
    double acc[1] = {0.0};
    line_tile<NP, 0, MODE, LDS, P, CT, VT>(a, lt % M.lgx, lt / M.lgx, M.lgx, stage, acc);
    if constexpr (MODE == LINE_BICG_S) {
      block_reduce_store<1>(acc, e.partial + (long)e.qslot * e.pstride, g, threadIdx.x, 64,
                            e.pstride);
      __syncthreads();
    }
  }
}

struct KEnsVec {
  const EnsMember* mem;
  const int* tmem;
  int ntiles, chunk, np1;
  const KState* st;
};

__device__ __forceinline__ void vec_range(const KEnsVec& e, const EnsMember& M, int g, long& e0,
                                          long& e1) {
  const long base = (long)e.np1 * M.off, len = (long)e.np1 * M.n;
  e0 = base + (long)(g - M.bt0) * e.chunk;
  e1 = min(base + len, e0 + (long)e.chunk);
}

template <typename VT>
__global__ __launch_bounds__(256) void k_ens_update(KEnsVec e, VT* __restrict__ x,
                                                    const VT* __restrict__ ph,
                                                    const VT* __restrict__ rh,
                                                    VT* __restrict__ r, const VT* __restrict__ t,
                                                    const VT* __restrict__ rt, double* partial,
                                                    long pstride) {
  const int tid = threadIdx.x;
  const int nb8 = gridDim.x >> 3, per = (e.ntiles + 7) >> 3;
  const int g0 = (blockIdx.x & 7) * per, g1 = min(e.ntiles, g0 + per);
  for (int g = g0 + (blockIdx.x >> 3); g < g1; g += nb8) {
    const int m = e.tmem[g];
    const KState* st = e.st + m;
    if (st->done) continue;
    const EnsMember M = e.mem[m];
    long a0, a1;
    vec_range(e, M, g, a0, a1);
    const double alpha = st->alpha, omega = st->omega;
    double acc[2] = {0.0, 0.0};
    if (st->half) {
      for (long q = a0 + tid; q < a1; q += 256) x[q] = (VT)((double)x[q] + alpha * (double)ph[q]);
    } else {
      for (long q = a0 + tid; q < a1; q += 256)
        bicg_upd_elem<VT>(q, x, ph, rh, r, t, rt, alpha, omega, acc);
    }
    block_reduce_store<2>(acc, partial, g, tid, 256, pstride);
    __syncthreads();
  }
}

// d = 0 for members still refining (start of an inner solve).
__global__ __launch_bounds__(256) void k_ens_zero(KEnsVec e, float* __restrict__ d) {
  const int nb8 = gridDim.x >> 3, per = (e.ntiles + 7) >> 3;
  const int g0 = (blockIdx.x & 7) * per, g1 = min(e.ntiles, g0 + per);
  for (int g = g0 + (blockIdx.x >> 3); g < g1; g += nb8) {
    const int m = e.tmem[g];
    if (e.st[m].outer_done) continue;
    const EnsMember M = e.mem[m];
    long a0, a1;
    vec_range(e, M, g, a0, a1);
    for (long q = a0 + threadIdx.x; q < a1; q += 256) d[q] = 0.0f;
  }
}

// x: members whose current iterate sits in buffer 1 of the pair get it copied to buffer 0.
__global__ __launch_bounds__(256) void k_ens_xcopy(KEnsVec e, double* __restrict__ xa,
                                                   const double* __restrict__ xb) {
  const int nb8 = gridDim.x >> 3, per = (e.ntiles + 7) >> 3;
  const int g0 = (blockIdx.x & 7) * per, g1 = min(e.ntiles, g0 + per);
  for (int g = g0 + (blockIdx.x >> 3); g < g1; g += nb8) {
    const int m = e.tmem[g];
    if (e.st[m].xpar == 0) continue;
    const EnsMember M = e.mem[m];
    long a0, a1;
    vec_range(e, M, g, a0, a1);
    for (long q = a0 + threadIdx.x; q < a1; q += 256) xa[q] = xb[q];
  }
}

// Per-member solve state at the start of a solve.
__global__ void k_ens_init(KState* st, int nmem, double tol, double inner_tol, int max_inner,
                           int* xflags) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m == 0) {
    xflags[0] = 0;  // the initial guess has a nonzero entry
    xflags[1] = 1;  // the initial guess is all zero
  }
  if (m >= nmem) return;
  KState* s = st + m;
  s->rho = 0.0; s->rho_old = 0.0; s->alpha = 0.0; s->omega = 0.0; s->beta = 0.0;
  s->rnorm0 = 0.0; s->rnorm = 0.0; s->tol = 0.0;
  s->iter = 0; s->maxiter = max_inner; s->done = 0; s->status = 0;
  s->half = 0; s->result = 0; s->first = 0; s->pad = 0;
  s->outer_r0 = 0.0; s->outer_r = 0.0; s->outer_tol = tol; s->inner_tol = inner_tol;
  s->outer = 0; s->inner_total = 0; s->outer_done = 0; s->pad2 = 0;
  s->xpar = 0; s->pad3 = 0; s->xnz = 0; s->xz = 0;
  for (int q = 0; q < 16; ++q) s->hist[q] = 0;
}

// Per-member state for timing iterations: active, first iteration, rho = 1, never exits.
__global__ void k_ens_tune_init(KState* st, int nmem) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmem) return;
  KState* s = st + m;
  s->rho = 1.0; s->rho_old = 1.0; s->alpha = 0.0; s->omega = 0.0; s->beta = 0.0;
  s->rnorm0 = 1.0; s->rnorm = 1.0; s->tol = 0.0;
  s->iter = 1; s->maxiter = 1 << 30; s->done = 0; s->status = 1;
  s->half = 0; s->result = 0; s->first = 0; s->pad = 0;
  s->outer_done = 0;
}

// One workgroup per member: sums the member's partials of each quantity in a fixed order
// (quantity q from producer (srcs >> 2q) & 3), then runs the scalar tail on its state.
__global__ __launch_bounds__(256) void k_ens_finalize(int kind, int nq, int srcs,
                                                      const double* partial, long pstride,
                                                      const EnsMember* mem, KState* st0) {
  __shared__ double sh[4][4];
  const int m = blockIdx.x, tid = threadIdx.x;
  const EnsMember M = mem[m];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q < nq) {
      const int src = (srcs >> (2 * q)) & 3;
      const int t0 = src == SRC_A ? M.at0 : (src == SRC_L ? M.lt0 : M.bt0);
      const int nt = src == SRC_A ? M.ant : (src == SRC_L ? M.lnt : M.bnt);
      const double* pq = partial + (long)q * pstride + t0;
      double s = 0.0;
      for (int b = tid; b < nt; b += 256) s += pq[b];
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
  if (tid != 0) return;
  double S[4];
  for (int q = 0; q < 4; ++q) S[q] = q < nq ? ((sh[q][0] + sh[q][1]) + sh[q][2]) + sh[q][3] : 0.0;
  fin_logic(kind, S, st0 + m, 0, 0, nq, nullptr);
}

}  // namespace asam

using namespace asam;

// ======================================================================== context
struct asam_ens_ctx {
  int device = 0, ncu = 256;
  hipStream_t stream = nullptr;
  int nmem = 0, np1 = 1, grad_full = 1;
  std::vector<asam_sound_grid> grids;
  std::vector<EnsMember> mh;
  std::vector<int> ready;
  long Ntot = 0, Ftot = 0, n = 0;
  int maxnz = 0;
  Coef<double> c64;
  Coef<float> c32;
  double* xb = nullptr;                  // buffer 1 of the refinement pair
  double *xh = nullptr, *bh = nullptr;   // device copies for host-array solves
  float *r = nullptr, *rt = nullptr, *p = nullptr, *ph = nullptr, *v = nullptr, *rh = nullptr,
        *t = nullptr, *d = nullptr;
  double* partial = nullptr;
  long pstride = 0, pcap = 0;
  int *tA = nullptr, *tL = nullptr, *tB = nullptr;
  int nA = 0, nL = 0, nB = 0;
  long capA = 0, capL = 0, capB = 0;
  EnsMember* md = nullptr;
  KState* st = nullptr;
  KState* sth = nullptr;  // pinned
  unsigned char* dead = nullptr;  // cells whose residual is 0, from the FP64 coefficients
  int* xflags = nullptr;          // initial guess: [0] has a nonzero entry, [1] is all zero
  asam_sound_control ctl{};
  asam_ens_tuning tune{};
  std::vector<void*> allocs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // captured schedule
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const double* g_x = nullptr;
  const double* g_b = nullptr;
  double g_tol = -1.0;
  std::vector<int> g_sched;
  size_t g_nodes = 0;
  std::vector<int> sched;  // inner iterations per refinement step for the next solve
  asam_ens_stats stats{};
  int eager = 0;
  std::string err;
};

#define EHIP(ctx, expr)                                               \
  do {                                                                \
    hipError_t e_ = (expr);                                           \
    if (e_ != hipSuccess) {                                           \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_); \
      return ASAM_E_HIP;                                              \
    }                                                                 \
  } while (0)

static int efail(asam_ens_ctx* c, int code, const char* msg) {
  if (c) c->err = msg;
  return code;
}

template <typename T>
static int ealloc(asam_ens_ctx* c, T** p, size_t count) {
  void* q = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&q, count * sizeof(T));
  if (e != hipSuccess) {
    c->err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return ASAM_E_NOMEM;
  }
  c->allocs.push_back(q);
  *p = (T*)q;
  return 0;
}

static void efree(asam_ens_ctx* c, void* p) {
  if (!p) return;
  for (size_t i = 0; i < c->allocs.size(); ++i)
    if (c->allocs[i] == p) {
      c->allocs.erase(c->allocs.begin() + i);
      break;
    }
  (void)hipFree(p);
}

static void drop_graph(asam_ens_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->g_sched.clear();
  c->g_nodes = 0;
}

static Geom member_geom(const asam_ens_ctx* c, int m) {
  Geom g;
  const asam_sound_grid& q = c->grids[m];
  g.nx = q.nx; g.ny = q.ny; g.nz = q.nz; g.np1 = q.np1;
  g.N = (long)q.nx * q.ny * q.nz;
  g.fa = q.fac_anela;
  g.tune = Tuning();
  g.tune.line_tiled = 1;
  return g;
}

template <typename T>
static Coef<T> member_coef(const Coef<T>& pool, const EnsMember& M, int np1) {
  Coef<T> k = pool;
  k.kx += M.off; k.ky += M.off; k.kz += M.off; k.vol += M.off; k.dut += M.off;
  if (np1 == 2) { k.tx += M.off; k.ty += M.off; k.tz += M.off; k.gz += M.off; }
  if (k.pkz) k.pkz += 4 * M.off;
  const long fo = (long)np1 * np1 * M.foff;
  k.pl += lower_words(np1) * M.foff; k.pt += fo; k.pw += fo;
  return k;
}

static int grid_for(int ntiles, int pf, int ncu) {
  long nb = ((long)ntiles + 7) / 8 * 8;
  if (pf > 0) {
    long cap = ((long)pf * ncu + 7) / 8 * 8;
    if (nb > cap) nb = cap;
  }
  if (nb < 8) nb = 8;
  return (int)nb;
}

static int valid_tuning(const asam_ens_tuning& t) {
  if (t.apply_bx != 64 && t.apply_bx != 128 && t.apply_bx != 256) return 0;
  if (t.apply_zsplit < 1 || t.apply_zsplit > 8) return 0;
  if (t.apply_rows < 0 || t.apply_rows > 2) return 0;
  if (t.apply_pf < 0 || t.apply_pf > 64 || t.line_pf < 0 || t.line_pf > 64) return 0;
  if (t.blas_pf < 0 || t.blas_pf > 64) return 0;
  if (t.line_window != 4 && t.line_window != 8) return 0;
  if (t.line_lds < 0 || t.line_lds > 160) return 0;
  if (t.blas_chunk < 1024 || t.blas_chunk > 65536 || (t.blas_chunk & 255)) return 0;
  return 1;
}

// Tile tables of the current tuning (device copies), partial-sum stride.
static int build_tables(asam_ens_ctx* c) {
  const int bx = c->tune.apply_bx, chunk = c->tune.blas_chunk;
  const int by = (256 / bx) * (c->tune.apply_rows == 2 ? 2 : 1);  // rows per apply tile
  long gA = 0, gL = 0, gB = 0;
  for (int m = 0; m < c->nmem; ++m) {
    EnsMember& M = c->mh[m];
    M.agx = (M.nx + bx - 1) / bx;
    M.agy = (M.ny + by - 1) / by;
    const int zs = std::min(c->tune.apply_zsplit, M.nz);
    M.akp = (M.nz + zs - 1) / zs;
    const int nzt = (M.nz + M.akp - 1) / M.akp;
    const long ant = (long)M.agx * M.agy * nzt;
    M.lgx = (M.nx + 63) / 64;
    const long lnt = (long)M.lgx * M.ny;
    const long len = (long)c->np1 * M.n;
    const long bnt = (len + chunk - 1) / chunk;
    if (gA + ant > INT_MAX / 2 || gL + lnt > INT_MAX / 2 || gB + bnt > INT_MAX / 2)
      return efail(c, ASAM_E_INVALID, "ensemble too large for the tile tables");
    M.ant = (int)ant; M.at0 = (int)gA; gA += ant;
    M.lnt = (int)lnt; M.lt0 = (int)gL; gL += lnt;
    M.bnt = (int)bnt; M.bt0 = (int)gB; gB += bnt;
  }
  std::vector<int> A(gA), L(gL), B(gB);
  for (int m = 0; m < c->nmem; ++m) {
    const EnsMember& M = c->mh[m];
    std::fill(A.begin() + M.at0, A.begin() + M.at0 + M.ant, m);
    std::fill(L.begin() + M.lt0, L.begin() + M.lt0 + M.lnt, m);
    std::fill(B.begin() + M.bt0, B.begin() + M.bt0 + M.bnt, m);
  }
  EHIP(c, hipStreamSynchronize(c->stream));  // no launch may still read the old tables
  auto grow = [&](int** p, long& cap, long need) -> int {
    if (need <= cap) return 0;
    efree(c, *p);
    *p = nullptr;
    cap = 0;
    int rc = ealloc(c, p, need);
    if (!rc) cap = need;
    return rc;
  };
  int rc = grow(&c->tA, c->capA, gA) | grow(&c->tL, c->capL, gL) | grow(&c->tB, c->capB, gB);
  if (rc) return ASAM_E_NOMEM;
  const long ps = std::max(gA, std::max(gL, gB));
  if (ps > c->pcap) {
    efree(c, c->partial);
    c->partial = nullptr;
    c->pcap = 0;
    if (ealloc(c, &c->partial, 4 * ps)) return ASAM_E_NOMEM;
    c->pcap = ps;
  }
  c->pstride = ps;
  c->nA = (int)gA; c->nL = (int)gL; c->nB = (int)gB;
  EHIP(c, hipMemcpy(c->tA, A.data(), gA * sizeof(int), hipMemcpyHostToDevice));
  EHIP(c, hipMemcpy(c->tL, L.data(), gL * sizeof(int), hipMemcpyHostToDevice));
  EHIP(c, hipMemcpy(c->tB, B.data(), gB * sizeof(int), hipMemcpyHostToDevice));
  EHIP(c, hipMemcpy(c->md, c->mh.data(), c->nmem * sizeof(EnsMember), hipMemcpyHostToDevice));
  drop_graph(c);
  return 0;
}

// -------------------------------------------------------------------- launches
namespace {

KEnsVec vec_args(const asam_ens_ctx* c) {
  KEnsVec e;
  e.mem = c->md; e.tmem = c->tB; e.ntiles = c->nB; e.chunk = c->tune.blas_chunk;
  e.np1 = c->np1; e.st = c->st;
  return e;
}

void fin(asam_ens_ctx* c, int kind, int nq, int srcs) {
  hipLaunchKernelGGL(k_ens_finalize, dim3(c->nmem), dim3(256), 0, c->stream, kind, nq, srcs,
                     (const double*)c->partial, c->pstride, (const EnsMember*)c->md, c->st);
}

template <int NP, int FL, int PAR, typename CT, typename VX, typename VO, typename VU>
static void ens_apply_go(asam_ens_ctx* c, dim3 grid, dim3 blk,
                         const KEnsApply<CT, VX, VO, VU>& e) {
  if (c->tune.apply_rows == 2)
    hipLaunchKernelGGL((k_ens_apply<NP, FL, PAR, 2, CT, VX, VO, VU>), grid, blk, 0, c->stream,
                       e);
  else
    hipLaunchKernelGGL((k_ens_apply<NP, FL, PAR, 1, CT, VX, VO, VU>), grid, blk, 0, c->stream,
                       e);
}

template <int NP>
void apply_outer(asam_ens_ctx* c, bool first, double* x, const double* b) {
  KEnsApply<double, double, float, float> e;
  e.c = c->c64; e.mem = c->md; e.tmem = c->tA; e.ntiles = c->nA; e.bxw = c->tune.apply_bx;
  e.xa = x; e.xb = c->xb; e.dx = c->d; e.b = b; e.y = c->r; e.y2 = c->rt; e.u = nullptr;
  e.partial = c->partial; e.pstride = c->pstride; e.st = c->st; e.skip_outer = 1;
  const dim3 grid(grid_for(c->nA, c->tune.apply_pf, c->ncu));
  const dim3 blk(c->tune.apply_bx, 256 / c->tune.apply_bx);
  if (first) {
    // a scan of the initial guess picks on the device which launch writes the residual:
    // the full FP64 apply, or for an all-zero guess the x = 0 form (b outside the dead
    // cells: the same values and partial sums from a quarter of the bytes)
    launch_any_nonzero(c->stream, c->n, x, c->xflags, c->xflags + 1);
    e.skip = c->xflags + 1;
    ens_apply_go<NP, AF_B | AF_Y2 | AF_YY, 0>(c, grid, blk, e);
    e.skip = c->xflags;
    e.dead = c->dead;
    ens_apply_go<NP, AF_X0 | AF_B | AF_Y2 | AF_YY, 0>(c, grid, blk, e);
  } else {
    ens_apply_go<NP, AF_B | AF_Y2 | AF_DX | AF_YY, 1>(c, grid, blk, e);
  }
}

template <int NP, int FL>
void apply_inner(asam_ens_ctx* c, float* x, float* y, const float* u) {
  KEnsApply<float, float, float, float> e;
  e.c = c->c32; e.mem = c->md; e.tmem = c->tA; e.ntiles = c->nA; e.bxw = c->tune.apply_bx;
  e.xa = x; e.xb = nullptr; e.dx = nullptr; e.b = nullptr; e.y = y; e.y2 = nullptr; e.u = u;
  e.partial = c->partial; e.pstride = c->pstride; e.st = c->st; e.skip_outer = 0;
  ens_apply_go<NP, FL, 0>(c, dim3(grid_for(c->nA, c->tune.apply_pf, c->ncu)),
                         dim3(c->tune.apply_bx, 256 / c->tune.apply_bx), e);
}

template <int NP, int MODE, int LDS, int P>
void line_go(asam_ens_ctx* c, const KEnsLine<float, float>& e, size_t lds) {
  hipLaunchKernelGGL((k_ens_line<NP, MODE, LDS, P, float, float>),
                     dim3(grid_for(c->nL, c->tune.line_pf, c->ncu)), dim3(64), lds, c->stream,
                     e);
}

template <int NP, int MODE>
void line_inner(asam_ens_ctx* c, const float* src, float* p, const float* v, float* out,
                int qslot) {
  KEnsLine<float, float> e;
  e.c = c->c32; e.mem = c->md; e.tmem = c->tL; e.ntiles = c->nL;
  e.src = src; e.p = p; e.v = v; e.out = out; e.st = c->st;
  e.partial = c->partial; e.pstride = c->pstride; e.qslot = qslot;
  e.kl = lds_levels(c->maxnz, NP, sizeof(float), lds_budget_for(c->tune.line_lds));
  const size_t lds = (size_t)e.kl * 64 * NP * sizeof(float);
  const bool use = e.kl > 0;
  if (use) {
    if (c->tune.line_window == 8) line_go<NP, MODE, 1, 8>(c, e, lds);
    else line_go<NP, MODE, 1, 4>(c, e, lds);
  } else {
    if (c->tune.line_window == 8) line_go<NP, MODE, 0, 8>(c, e, 0);
    else line_go<NP, MODE, 0, 4>(c, e, 0);
  }
}

void update_inner(asam_ens_ctx* c) {
  hipLaunchKernelGGL(k_ens_update<float>, dim3(grid_for(c->nB, c->tune.blas_pf, c->ncu)),
                     dim3(256), 0, c->stream, vec_args(c), c->d, (const float*)c->ph,
                     (const float*)c->rh, c->r, (const float*)c->t, (const float*)c->rt,
                     c->partial, c->pstride);
}

constexpr int srcs(int a, int b = 0, int q = 0, int d = 0) {
  return a | (b << 2) | (q << 4) | (d << 6);
}

// One FP32 BiCGStab iteration of every member (the single-domain bicg_iteration sequence).
template <int NP>
void inner_iteration(asam_ens_ctx* c) {
  line_inner<NP, LINE_BICG_P>(c, c->r, c->p, c->v, c->ph, 0);
  apply_inner<NP, AF_YU>(c, c->ph, c->v, c->rt);
  fin(c, FIN_BICG_ALPHA, 2, srcs(SRC_A, SRC_A));
  line_inner<NP, LINE_BICG_S>(c, nullptr, c->r, c->v, c->rh, 2);
  apply_inner<NP, AF_YY | AF_YU>(c, c->rh, c->t, c->r);
  fin(c, FIN_BICG_HALF_OMEGA, 3, srcs(SRC_A, SRC_A, SRC_L));
  update_inner(c);
  fin(c, FIN_BICG_END, 2, srcs(SRC_B, SRC_B));
}

template <int NP>
void outer_step(asam_ens_ctx* c, bool first, double* x, const double* b) {
  apply_outer<NP>(c, first, x, b);
  fin(c, FIN_MIXED_OUTER, 1, srcs(SRC_A));
}

void zero_d(asam_ens_ctx* c) {
  hipLaunchKernelGGL(k_ens_zero, dim3(grid_for(c->nB, c->tune.blas_pf, c->ncu)), dim3(256), 0,
                     c->stream, vec_args(c), c->d);
}

void xcopy(asam_ens_ctx* c, double* x) {
  hipLaunchKernelGGL(k_ens_xcopy, dim3(grid_for(c->nB, c->tune.blas_pf, c->ncu)), dim3(256), 0,
                     c->stream, vec_args(c), x, (const double*)c->xb);
}

// The whole schedule of one solve: init, refinement step 0, then per entry of sched an
// inner solve of sched[q] iterations followed by the next refinement step; last, the
// iterate is gathered into x.
template <int NP>
void enqueue_schedule(asam_ens_ctx* c, double* x, const double* b, double tol,
                      const std::vector<int>& sched) {
  hipLaunchKernelGGL(k_ens_init, dim3((c->nmem + 63) / 64), dim3(64), 0, c->stream, c->st,
                     c->nmem, tol, c->ctl.inner_tol, c->ctl.max_inner, c->xflags);
  outer_step<NP>(c, true, x, b);
  for (size_t q = 0; q < sched.size(); ++q) {
    zero_d(c);
    for (int i = 0; i < sched[q]; ++i) inner_iteration<NP>(c);
    outer_step<NP>(c, false, x, b);
  }
  xcopy(c, x);
}

int read_states(asam_ens_ctx* c) {
  EHIP(c, hipMemcpyAsync(c->sth, c->st, sizeof(KState) * c->nmem, hipMemcpyDeviceToHost,
                         c->stream));
  EHIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

template <int NP>
int solve_np(asam_ens_ctx* c, double* x, const double* b, double tol) {
  std::vector<int> sched = c->sched;
  if ((int)sched.size() > c->ctl.max_outer) sched.resize(c->ctl.max_outer);
  for (int& s : sched) s = std::min(s, c->ctl.max_inner);
  c->stats = asam_ens_stats{};
  c->eager = 0;
  EHIP(c, hipEventRecord(c->ev0, c->stream));
  if (c->tune.use_graph) {
    if (!c->gexec || c->g_x != x || c->g_b != b || c->g_tol != tol || c->g_sched != sched) {
      drop_graph(c);
      EHIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      enqueue_schedule<NP>(c, x, b, tol, sched);
      hipGraph_t g = nullptr;
      EHIP(c, hipStreamEndCapture(c->stream, &g));
      c->graph = g;
      EHIP(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
      size_t nn = 0;
      EHIP(c, hipGraphGetNodes(c->graph, nullptr, &nn));
      c->g_nodes = nn;
      c->g_x = x; c->g_b = b; c->g_tol = tol; c->g_sched = sched;
    }
    EHIP(c, hipGraphLaunch(c->gexec, c->stream));
    c->stats.graph_launches = 1;
    c->stats.graph_nodes = (int)c->g_nodes;
  } else {
    enqueue_schedule<NP>(c, x, b, tol, sched);
    c->eager += 1 + 2 + (int)sched.size() * 3 + 1;
    for (int s : sched) c->eager += 8 * s;
  }
  int rc = read_states(c);
  if (rc) return rc;
  // Members the schedule did not finish keep refining in eagerly enqueued steps.
  int steps = (int)sched.size();
  auto all_outer = [&] {
    for (int m = 0; m < c->nmem; ++m)
      if (!c->sth[m].outer_done) return false;
    return true;
  };
  bool continued = false;
  while (!all_outer() && steps < c->ctl.max_outer) {
    continued = true;
    zero_d(c);
    c->eager++;
    int issued = 0, chunk = 4;
    while (issued < c->ctl.max_inner) {
      const int q = std::min(chunk, c->ctl.max_inner - issued);
      for (int i = 0; i < q; ++i) inner_iteration<NP>(c);
      c->eager += 8 * q;
      issued += q;
      rc = read_states(c);
      if (rc) return rc;
      bool all = true;
      for (int m = 0; m < c->nmem; ++m)
        if (!c->sth[m].outer_done && !c->sth[m].done) all = false;
      if (all) break;
      chunk = 2;
    }
    outer_step<NP>(c, false, x, b);
    c->eager += 2;
    steps++;
    rc = read_states(c);
    if (rc) return rc;
  }
  if (continued) {
    xcopy(c, x);
    c->eager++;
  }
  EHIP(c, hipEventRecord(c->ev1, c->stream));
  EHIP(c, hipEventSynchronize(c->ev1));
  EHIP(c, hipGetLastError());
  float ms = 0.f;
  EHIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->stats.ms = ms;
  c->stats.eager_launches = c->eager;
  // Next schedule: per refinement step the most inner iterations any member took, + 1.
  int outer_max = 0;
  for (int m = 0; m < c->nmem; ++m) outer_max = std::max(outer_max, c->sth[m].outer);
  c->stats.outer_steps = outer_max;
  std::vector<int> next;
  for (int q = 0; q < std::min(outer_max, 16); ++q) {
    int mx = 0;
    for (int m = 0; m < c->nmem; ++m)
      if (c->sth[m].outer > q) mx = std::max(mx, c->sth[m].hist[q]);
    next.push_back(std::max(1, mx + 1));
  }
  if (!next.empty()) c->sched = next;
  return 0;
}

}  // namespace

// ======================================================================== C-ABI
extern "C" int asam_ens_create(asam_ens_ctx** out, int nmem, const asam_sound_grid* grids) {
  if (!out || !grids || nmem < 1) return ASAM_E_INVALID;
  *out = nullptr;
  const int np1 = grids[0].np1;
  if (np1 != 1 && np1 != 2) return ASAM_E_INVALID;
  long Ntot = 0, Ftot = 0;
  int maxnz = 0;
  for (int m = 0; m < nmem; ++m) {
    const asam_sound_grid& q = grids[m];
    if (q.nx < 1 || q.ny < 1 || q.nz < 1 || q.np1 != np1 || q.grad_full != grids[0].grad_full)
      return ASAM_E_INVALID;
    Geom g;
    g.nx = q.nx; g.ny = q.ny; g.nz = q.nz; g.np1 = np1; g.N = (long)q.nx * q.ny * q.nz;
    g.tune = Tuning();
    g.tune.line_tiled = 1;
    Ntot += g.N;
    Ftot += fac_cells(g);
    maxnz = std::max(maxnz, q.nz);
  }
  if (Ntot > (1L << 33)) return ASAM_E_INVALID;
  asam_ens_ctx* c = new asam_ens_ctx();
  *out = c;
  c->nmem = nmem; c->np1 = np1; c->grad_full = grids[0].grad_full;
  c->grids.assign(grids, grids + nmem);
  c->Ntot = Ntot; c->Ftot = Ftot; c->n = Ntot * np1; c->maxnz = maxnz;
  c->ready.assign(nmem, 0);
  c->mh.resize(nmem);
  long off = 0, foff = 0;
  for (int m = 0; m < nmem; ++m) {
    EnsMember& M = c->mh[m];
    memset(&M, 0, sizeof M);
    const Geom g = member_geom(c, m);
    M.nx = g.nx; M.ny = g.ny; M.nz = g.nz; M.np1 = np1;
    M.off = off; M.n = g.N; M.foff = foff; M.fa = g.fa;
    off += g.N;
    foff += fac_cells(g);
  }
  asam_sound_default_control(&c->ctl);
  c->ctl.precision = ASAM_PREC_MIXED;
  c->tune.apply_bx = 256; c->tune.apply_zsplit = 1; c->tune.apply_pf = 0;
  c->tune.apply_rows = 2;
  c->tune.line_window = 8; c->tune.line_lds = 1; c->tune.line_pf = 0;
  c->tune.blas_chunk = 4096; c->tune.blas_pf = 0; c->tune.use_graph = 1;
  c->sched = {10, 6, 4};
  if (hipGetDevice(&c->device) != hipSuccess) return efail(c, ASAM_E_HIP, "no HIP device");
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) ==
          hipSuccess && ncu > 0)
    c->ncu = ncu;
  EHIP(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  EHIP(c, hipEventCreate(&c->ev0));
  EHIP(c, hipEventCreate(&c->ev1));
  (void)lds_budget();  // device query cached here, never during a graph capture
  const long N = Ntot, P = (long)np1 * np1 * Ftot, PL = (long)lower_words(np1) * Ftot;
  int rc = 0;
  auto coef = [&](auto& k) {
    rc |= ealloc(c, &k.kx, N); rc |= ealloc(c, &k.ky, N); rc |= ealloc(c, &k.kz, N);
    if (np1 == 2) {
      rc |= ealloc(c, &k.tx, N); rc |= ealloc(c, &k.ty, N); rc |= ealloc(c, &k.tz, N);
      rc |= ealloc(c, &k.gz, N);
    }
    rc |= ealloc(c, &k.vol, N); rc |= ealloc(c, &k.dut, N);
    rc |= ealloc(c, &k.pl, PL); rc |= ealloc(c, &k.pt, P); rc |= ealloc(c, &k.pw, P);
  };
  coef(c->c64);
  coef(c->c32);
  if (np1 == 2) rc |= ealloc(c, &c->c32.pkz, 4 * N);  // packed (kz, tz, gz, vol), FP32 apply
  rc |= ealloc(c, &c->xb, c->n);
  float** v32[] = {&c->r, &c->rt, &c->p, &c->ph, &c->v, &c->rh, &c->t, &c->d};
  for (float** q : v32) rc |= ealloc(c, q, c->n);
  rc |= ealloc(c, &c->md, nmem);
  rc |= ealloc(c, &c->st, nmem);
  rc |= ealloc(c, &c->dead, c->Ntot);
  rc |= ealloc(c, &c->xflags, 2);
  if (rc) return ASAM_E_NOMEM;
  EHIP(c, hipHostMalloc((void**)&c->sth, sizeof(KState) * nmem, hipHostMallocDefault));
  memset(c->sth, 0, sizeof(KState) * nmem);
  // FP32 vectors start finite (the autotuner times kernels on them before any solve)
  for (float** q : v32) launch_fill(c->stream, c->n, 1, *q, 0.0);
  EHIP(c, hipGetLastError());
  return build_tables(c);
}

extern "C" int asam_ens_destroy(asam_ens_ctx* c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  for (void* p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  if (c->sth) (void)hipHostFree(c->sth);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

extern "C" const char* asam_ens_last_error(const asam_ens_ctx* c) {
  return c ? c->err.c_str() : "null context";
}

extern "C" long asam_ens_length(const asam_ens_ctx* c) { return c ? c->n : 0; }

extern "C" int asam_ens_member_range(const asam_ens_ctx* c, int m, long* offset, long* length) {
  if (!c || m < 0 || m >= c->nmem || !offset || !length) return ASAM_E_INVALID;
  *offset = (long)c->np1 * c->mh[m].off;
  *length = (long)c->np1 * c->mh[m].n;
  return 0;
}

extern "C" int asam_ens_set_control(asam_ens_ctx* c, const asam_sound_control* ctl) {
  if (!c || !ctl) return ASAM_E_INVALID;
  if (ctl->method != ASAM_METHOD_BICGSTAB || ctl->precision != ASAM_PREC_MIXED)
    return efail(c, ASAM_E_INVALID, "ensembles run the mixed-precision BiCGStab solve");
  if (ctl->max_inner < 1 || ctl->max_outer < 1 || !(ctl->inner_tol > 0.0))
    return efail(c, ASAM_E_INVALID, "max_inner, max_outer >= 1 and inner_tol > 0");
  c->ctl = *ctl;
  drop_graph(c);
  return 0;
}

extern "C" int asam_ens_setup(asam_ens_ctx* c, int m, const double* FU, const double* FV,
                              const double* FW, const double* V, const double* dx,
                              const double* dy, const double* dz, const double* rho,
                              const double* rth, const double* T) {
  if (!c) return ASAM_E_INVALID;
  if (m < 0 || m >= c->nmem) return efail(c, ASAM_E_INVALID, "member index");
  if (!FU || !FV || !FW || !V || !dx || !dy || !dz || !rho || !rth || !T)
    return efail(c, ASAM_E_INVALID, "null input");
  EHIP(c, hipSetDevice(c->device));
  const Geom g = member_geom(c, m);
  const long N = g.N;
  for (long q = 0; q < N; ++q)
    if (!(V[q] >= 0.0) || !isfinite(V[q])) return efail(c, ASAM_E_NAN, "VolC must be finite, >= 0");
  const long nfu = (long)(g.nx + 1) * g.ny * g.nz, nfv = (long)g.nx * (g.ny + 1) * g.nz,
             nfw = (long)g.nx * g.ny * (g.nz + 1);
  double *dFU, *dFV, *dFW, *dV, *ddx, *ddy, *ddz, *drho, *drth, *dT, *tmp, *th;
  int rc = 0;
  rc |= ealloc(c, &dFU, nfu); rc |= ealloc(c, &dFV, nfv); rc |= ealloc(c, &dFW, nfw);
  rc |= ealloc(c, &dV, N); rc |= ealloc(c, &ddx, g.nx); rc |= ealloc(c, &ddy, g.ny);
  rc |= ealloc(c, &ddz, g.nz); rc |= ealloc(c, &drho, N); rc |= ealloc(c, &drth, N);
  rc |= ealloc(c, &dT, N); rc |= ealloc(c, &tmp, N); rc |= ealloc(c, &th, nfu + nfv + nfw);
  if (rc) return ASAM_E_NOMEM;
  hipStream_t s = c->stream;
  EHIP(c, hipMemcpyAsync(dFU, FU, nfu * 8, hipMemcpyHostToDevice, s));
  EHIP(c, hipMemcpyAsync(dFV, FV, nfv * 8, hipMemcpyHostToDevice, s));
  EHIP(c, hipMemcpyAsync(dFW, FW, nfw * 8, hipMemcpyHostToDevice, s));
  EHIP(c, hipMemcpyAsync(dV, V, N * 8, hipMemcpyHostToDevice, s));
  EHIP(c, hipMemcpyAsync(ddx, dx, g.nx * 8, hipMemcpyHostToDevice, s));
  EHIP(c, hipMemcpyAsync(ddy, dy, g.ny * 8, hipMemcpyHostToDevice, s));
  EHIP(c, hipMemcpyAsync(ddz, dz, g.nz * 8, hipMemcpyHostToDevice, s));
  EHIP(c, hipMemcpyAsync(drho, rho, N * 8, hipMemcpyHostToDevice, s));
  EHIP(c, hipMemcpyAsync(drth, rth, N * 8, hipMemcpyHostToDevice, s));
  EHIP(c, hipMemcpyAsync(dT, T, N * 8, hipMemcpyHostToDevice, s));
  const EnsMember& M = c->mh[m];
  Coef<double> k64 = member_coef(c->c64, M, c->np1);
  Coef<float> k32 = member_coef(c->c32, M, c->np1);
  const asam_sound_grid& q = c->grids[m];
  launch_weights(s, g, drho, drth, dT, nullptr, nullptr, nullptr, tmp, k64.dut);
  launch_face_theta(s, g, dV, tmp, th, th + nfu, th + nfu + nfv);
  launch_coef(s, g, q.grad_full, q.beta0_dt, q.grav, dFU, dFV, dFW, dV, ddx, ddy, ddz, th,
              th + nfu, th + nfu + nfv, k64);
  launch_factor(s, g, k64);
  launch_convert(s, g, k64, k32);
  launch_dead_mask(s, g, k64, c->dead + M.off);
  EHIP(c, hipGetLastError());
  EHIP(c, hipStreamSynchronize(s));
  double* tmps[] = {dFU, dFV, dFW, dV, ddx, ddy, ddz, drho, drth, dT, tmp, th};
  for (double* p : tmps) efree(c, p);
  c->ready[m] = 1;
  return 0;
}

static int check_ready(asam_ens_ctx* c) {
  for (int m = 0; m < c->nmem; ++m)
    if (!c->ready[m]) return efail(c, ASAM_E_STATE, "every member must be set up first");
  return 0;
}

extern "C" int asam_ens_solve_device(asam_ens_ctx* c, double* x, const double* b, double tol,
                                     asam_sound_info* info) {
  if (!c) return ASAM_E_INVALID;
  if (!x || !b) return efail(c, ASAM_E_INVALID, "null argument");
  if (!(tol >= 0.0)) return efail(c, ASAM_E_INVALID, "tol");
  int rc = check_ready(c);
  if (rc) return rc;
  EHIP(c, hipSetDevice(c->device));
  rc = c->np1 == 1 ? solve_np<1>(c, x, b, tol) : solve_np<2>(c, x, b, tol);
  if (rc) return rc;
  bool finite = true;
  for (int m = 0; m < c->nmem; ++m) {
    const KState& s = c->sth[m];
    if (!isfinite(s.outer_r)) finite = false;
    if (info) {
      asam_sound_info& o = info[m];
      memset(&o, 0, sizeof o);
      o.iterations = s.inner_total;
      o.restarts = s.outer;
      o.status = s.outer_done ? ASAM_CONVERGED : ASAM_NOT_CONVERGED;
      o.abs_residual = s.outer_r;
      o.rel_residual = s.outer_r0 > 0 ? s.outer_r / s.outer_r0 : 0.0;
      o.ms = c->stats.ms;
    }
  }
  if (!finite) return efail(c, ASAM_E_NAN, "non-finite residual");
  return 0;
}

extern "C" int asam_ens_solve(asam_ens_ctx* c, double* x, const double* b, double tol,
                              asam_sound_info* info) {
  if (!c) return ASAM_E_INVALID;
  if (!x || !b) return efail(c, ASAM_E_INVALID, "null argument");
  EHIP(c, hipSetDevice(c->device));
  if (!c->xh) {
    if (ealloc(c, &c->xh, c->n) || ealloc(c, &c->bh, c->n)) return ASAM_E_NOMEM;
  }
  EHIP(c, hipMemcpyAsync(c->xh, x, c->n * 8, hipMemcpyHostToDevice, c->stream));
  EHIP(c, hipMemcpyAsync(c->bh, b, c->n * 8, hipMemcpyHostToDevice, c->stream));
  int rc = asam_ens_solve_device(c, c->xh, c->bh, tol, info);
  if (rc && rc != ASAM_E_NAN) return rc;
  EHIP(c, hipMemcpyAsync(x, c->xh, c->n * 8, hipMemcpyDeviceToHost, c->stream));
  EHIP(c, hipStreamSynchronize(c->stream));
  return rc;
}

extern "C" int asam_ens_get_tuning(const asam_ens_ctx* c, asam_ens_tuning* t) {
  if (!c || !t) return ASAM_E_INVALID;
  *t = c->tune;
  return 0;
}

extern "C" int asam_ens_set_tuning(asam_ens_ctx* c, const asam_ens_tuning* t) {
  if (!c || !t) return ASAM_E_INVALID;
  if (!valid_tuning(*t)) return efail(c, ASAM_E_INVALID, "tuning out of range");
  EHIP(c, hipSetDevice(c->device));
  c->tune = *t;
  return build_tables(c);
}

extern "C" int asam_ens_get_stats(const asam_ens_ctx* c, asam_ens_stats* s) {
  if (!c || !s) return ASAM_E_INVALID;
  *s = c->stats;
  return 0;
}

extern "C" int asam_ens_dev_alloc(asam_ens_ctx* c, void** ptr, size_t bytes) {
  if (!c || !ptr) return ASAM_E_INVALID;
  EHIP(c, hipSetDevice(c->device));
  EHIP(c, hipMalloc(ptr, bytes ? bytes : 1));
  return 0;
}
extern "C" int asam_ens_dev_free(asam_ens_ctx* c, void* ptr) {
  if (!c) return ASAM_E_INVALID;
  if (ptr) {
    EHIP(c, hipStreamSynchronize(c->stream));
    EHIP(c, hipFree(ptr));
  }
  return 0;
}
extern "C" int asam_ens_dev_upload(asam_ens_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || !dst || !src) return ASAM_E_INVALID;
  EHIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  EHIP(c, hipStreamSynchronize(c->stream));
  return 0;
}
extern "C" int asam_ens_dev_download(asam_ens_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || !dst || !src) return ASAM_E_INVALID;
  EHIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  EHIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int asam_ens_dev_memset(asam_ens_ctx* c, void* dst, int value, size_t bytes) {
  if (!c || !dst) return ASAM_E_INVALID;
  EHIP(c, hipMemsetAsync(dst, value, bytes, c->stream));
  EHIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

// ------------------------------------------------------------------ autotuner
namespace {

template <typename F>
int time_launches(asam_ens_ctx* c, int reps, F&& launch, double* ms) {
  launch();  // warm-up
  EHIP(c, hipEventRecord(c->ev0, c->stream));
  for (int i = 0; i < reps; ++i) launch();
  EHIP(c, hipEventRecord(c->ev1, c->stream));
  EHIP(c, hipEventSynchronize(c->ev1));
  EHIP(c, hipGetLastError());
  float t = 0.f;
  EHIP(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
  *ms = t / reps;
  return 0;
}

// Objective: device time of whole FP32 inner BiCGStab iterations over every member
// (all kernels and finalizes in solve order), so each kernel runs in the cache context it
// has in a solve.  State: every member active, r = r~ = 1, x = 0, rho = 1, tol = 0.
template <int NP>
int time_inner(asam_ens_ctx* c, int iters, double* ms) {
  launch_fill(c->stream, c->n, 1, c->r, 1.0);
  launch_fill(c->stream, c->n, 1, c->rt, 1.0);
  launch_fill(c->stream, c->n, 1, c->d, 0.0);
  hipLaunchKernelGGL(k_ens_tune_init, dim3((c->nmem + 63) / 64), dim3(64), 0, c->stream, c->st,
                     c->nmem);
  EHIP(c, hipEventRecord(c->ev0, c->stream));
  for (int i = 0; i < iters; ++i) inner_iteration<NP>(c);
  EHIP(c, hipEventRecord(c->ev1, c->stream));
  EHIP(c, hipEventSynchronize(c->ev1));
  EHIP(c, hipGetLastError());
  float t = 0.f;
  EHIP(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
  *ms = (double)t / iters;
  return 0;
}

// Coordinate search over the launch shapes: apply, then line, then update.
template <int NP>
int autotune_np(asam_ens_ctx* c, int reps, double* best_ms) {
  auto measure = [&](double* out) -> int {
    double m = 0.0;
    int rc = build_tables(c);
    if (rc) return rc;
    rc = time_inner<NP>(c, 1, &m);  // warm
    if (rc) return rc;
    return time_inner<NP>(c, reps, out);
  };
  asam_ens_tuning bt = c->tune;
  double bms = 1e30;
  int rc;
  {
    const int bxs[] = {64, 128, 256}, zss[] = {1, 2, 4}, pfs[] = {0, 2, 8}, rbs[] = {1, 2};
    const asam_ens_tuning t0 = bt;
    for (int rb : rbs)
    for (int bx : bxs)
      for (int zs : zss)
        for (int pf : pfs) {
          c->tune = t0;
          c->tune.apply_bx = bx; c->tune.apply_zsplit = zs; c->tune.apply_pf = pf;
          c->tune.apply_rows = rb;
          double ms = 0;
          if ((rc = measure(&ms))) return rc;
          if (ms < bms) { bms = ms; bt = c->tune; }
        }
  }
  {
    const int ws[] = {4, 8}, ls[] = {0, 1}, pfs[] = {0, 4};
    const asam_ens_tuning t0 = bt;
    for (int w : ws)
      for (int l : ls)
        for (int pf : pfs) {
          c->tune = t0;
          c->tune.line_window = w; c->tune.line_lds = l; c->tune.line_pf = pf;
          double ms = 0;
          if ((rc = measure(&ms))) return rc;
          if (ms < bms) { bms = ms; bt = c->tune; }
        }
  }
  {
    const int chs[] = {2048, 8192, 32768}, pfs[] = {0, 4};
    const asam_ens_tuning t0 = bt;
    for (int ch : chs)
      for (int pf : pfs) {
        c->tune = t0;
        c->tune.blas_chunk = ch; c->tune.blas_pf = pf;
        double ms = 0;
        if ((rc = measure(&ms))) return rc;
        if (ms < bms) { bms = ms; bt = c->tune; }
      }
  }
  c->tune = bt;
  best_ms[0] = bms;
  best_ms[1] = best_ms[2] = 0.0;
  return build_tables(c);
}

}  // namespace

extern "C" int asam_ens_autotune(asam_ens_ctx* c, int reps, asam_ens_tuning* best, double* ms) {
  if (!c) return ASAM_E_INVALID;
  if (reps < 1 || reps > 1000) return efail(c, ASAM_E_INVALID, "reps");
  int rc = check_ready(c);
  if (rc) return rc;
  EHIP(c, hipSetDevice(c->device));
  double bm[3] = {0, 0, 0};
  rc = c->np1 == 1 ? autotune_np<1>(c, reps, bm) : autotune_np<2>(c, reps, bm);
  if (rc) return rc;
  if (best) *best = c->tune;
  if (ms) for (int i = 0; i < 3; ++i) ms[i] = bm[i];
  return 0;
}
