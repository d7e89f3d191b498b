// solver.hip -- C-ABI of the gfx950 sound stage solver: context, set-up, the
// SolveSound drop-in (BiCGStab / FGMRES / GMRES, FP64 or FP32-inner with FP64
// refinement), the adjoint pair, and device helpers.
//
// Krylov control flow stays on the device: every kernel of an iteration reads
// the KState flags and returns at once after convergence, so the host enqueues
// iterations in chunks and synchronises only between chunks; the iterates are
// exactly those of a loop that stops at the first converged iteration.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "asam_sound.h"
#include "internal.h"

using namespace asam;

struct asam_sound_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  asam_sound_grid grid{};
  asam_sound_control ctl{};
  Geom g{};
  int ready = 0;
  long n = 0;  // np1 * N
  Coef<double> c64;
  Coef<float> c32;
  double* thfaces = nullptr;  // face theta, FU|FV|FW shapes concatenated
  double *fu = nullptr, *fv = nullptr, *fw = nullptr;       // face areas (set-up input)
  double *dxv = nullptr, *dyv = nullptr, *dzv = nullptr;    // cell spacings
  double* pj_scratch = nullptr;  // host-array rhs/update staging (allocated on first use)
  double* partial = nullptr;
  unsigned int* tail_ctr = nullptr;  // arrival counter of the fused tails (zero at rest)
  unsigned char* dead = nullptr;  // cells whose residual is 0 (from the FP64 coefficients)
  KState* st = nullptr;
  KState* st_host = nullptr;  // pinned
  // FP64 work vectors
  double *r = nullptr, *rt = nullptr, *p = nullptr, *ph = nullptr, *v = nullptr,
         *rh = nullptr, *t = nullptr, *xd = nullptr, *bd = nullptr, *xalt = nullptr;
  // FP32 work vectors (mixed)
  float *r32 = nullptr, *rt32 = nullptr, *p32 = nullptr, *ph32 = nullptr, *v32 = nullptr,
        *rh32 = nullptr, *t32 = nullptr, *d32 = nullptr;
  // (F)GMRES basis
  int gm_m = 0;
  std::vector<double*> gv, gz;
  int launch_err = 0;  // an apply found no kernel for its option set (nothing launched)
  int gm32_m = 0;  // FP32 (F)GMRES basis of the mixed solve
  std::vector<float*> gv32, gz32;
  double* gtmp = nullptr;
  std::vector<void*> allocs;
  int counts[8] = {0};
  int last_iters = 4;     // BiCGStab iterations of the last plain solve (first chunk size)
  int inner_hist[32] = {0};  // inner iterations per refinement step of the last mixed solve
  int mixed_steps = 0;       // refinement steps of the last converged mixed solve (0: unknown)
  int mixed_cap = 0;         // inner iteration cap that solve ran with
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  // live per-kernel timing (asam_sound_profile_enable)
  struct ProfRec {
    int cls, tag, iter;
    double bytes, half_bytes;
    hipEvent_t e0, e1;
  };
  int prof_on = 0;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_next = 0;
  std::vector<ProfRec> prof_pending;
  double prof_ms[8] = {0}, prof_bytes[8] = {0};
  int prof_launches[8] = {0};
};

#define HIPCHK(ctx, expr)                                                         \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);             \
      return ASAM_E_HIP;                                                          \
    }                                                                             \
  } while (0)

static int fail(asam_sound_ctx* c, int code, const char* msg) {
  if (c) c->err = msg;
  return code;
}

template <typename T>
static int dalloc(asam_sound_ctx* c, T** p, size_t count) {
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

static void dfree(asam_sound_ctx* c, void* p) {
  if (!p) return;
  for (size_t i = 0; i < c->allocs.size(); ++i)
    if (c->allocs[i] == p) {
      c->allocs.erase(c->allocs.begin() + i);
      break;
    }
  (void)hipFree(p);
}

extern "C" int asam_sound_default_control(asam_sound_control* ctl) {
  if (!ctl) return ASAM_E_INVALID;
  ctl->method = ASAM_METHOD_BICGSTAB;
  ctl->step_gmres = 10;
  ctl->precision = ASAM_PREC_FP64;
  ctl->inner_tol = 1.0e-5;
  ctl->max_inner = 200;
  ctl->max_outer = 12;
  return 0;
}

extern "C" int asam_sound_get_tuning(const asam_sound_ctx* c, asam_sound_tuning* t) {
  if (!c || !t) return ASAM_E_INVALID;
  t->apply_variant = c->g.tune.apply_variant;
  t->apply_wg = c->g.tune.apply_wg;
  t->line_window = c->g.tune.line_window;
  t->line_lds = c->g.tune.line_lds;
  t->fused_tail = c->g.tune.fused_tail;
  t->apply_xcd = c->g.tune.apply_xcd;
  t->apply_bx = c->g.tune.apply_bx;
  t->line_tiled = c->g.tune.line_tiled;
  t->apply_band_rows = c->g.tune.apply_band_rows;
  t->apply_rows = c->g.tune.apply_rows;
  t->line_nt = c->g.tune.line_nt;
  t->apply_depth = c->g.tune.apply_depth;
  return 0;
}

extern "C" int asam_sound_set_tuning(asam_sound_ctx* c, const asam_sound_tuning* t) {
  if (!c || !t) return ASAM_E_INVALID;
  if (t->apply_variant < 0 || t->apply_variant > 3) return fail(c, ASAM_E_INVALID, "apply_variant");
  if (t->apply_rows != 2 && t->apply_rows != 4) return fail(c, ASAM_E_INVALID, "apply_rows");
  if (t->line_nt < 0 || t->line_nt > 2) return fail(c, ASAM_E_INVALID, "line_nt");
  if (t->apply_depth != 1 && t->apply_depth != 2) return fail(c, ASAM_E_INVALID, "apply_depth");
  if (t->apply_band_rows < 0) return fail(c, ASAM_E_INVALID, "apply_band_rows");
  if (t->apply_wg < 256 || t->apply_wg > 16384) return fail(c, ASAM_E_INVALID, "apply_wg");
  if (t->line_window != 0 && t->line_window != 1 && t->line_window != 4 && t->line_window != 8)
    return fail(c, ASAM_E_INVALID, "line_window");
  Tuning nt;
  nt.apply_variant = t->apply_variant;
  nt.apply_wg = t->apply_wg;
  nt.line_window = t->line_window;
  if (t->line_lds < 0 || t->line_lds > 160) return fail(c, ASAM_E_INVALID, "line_lds");
  nt.line_lds = t->line_lds;
  if (t->fused_tail < 0 || t->fused_tail > 2) return fail(c, ASAM_E_INVALID, "fused_tail");
  nt.fused_tail = t->fused_tail;
  if (t->apply_bx != 32 && t->apply_bx != 64 && t->apply_bx != 128 && t->apply_bx != 256)
    return fail(c, ASAM_E_INVALID, "apply_bx");
  nt.apply_xcd = t->apply_xcd ? 1 : 0;
  nt.apply_bx = t->apply_bx;
  nt.line_tiled = t->line_tiled ? 1 : 0;
  nt.apply_band_rows = t->apply_band_rows;
  nt.apply_rows = t->apply_rows;
  nt.line_nt = t->line_nt;
  nt.apply_depth = t->apply_depth;
  Geom tg = c->g;
  tg.tune = nt;
  if (apply_blocks(tg, 0) > kRedBlocks || apply_blocks(tg, 1) > kRedBlocks)
    return fail(c, ASAM_E_INVALID, "apply_wg too large");
  const bool relayout = nt.line_tiled != c->g.tune.line_tiled;
  c->g.tune = nt;
  if (relayout && c->ready) {  // the line factors live in the layout the tuning names
    HIPCHK(c, hipSetDevice(c->device));
    launch_factor(c->stream, c->g, c->c64);
    launch_convert(c->stream, c->g, c->c64, c->c32);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return 0;
}

extern "C" const char* asam_sound_last_error(const asam_sound_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int asam_sound_create(asam_sound_ctx** out, const asam_sound_grid* grid) {
  if (!out || !grid) return ASAM_E_INVALID;
  *out = nullptr;
  if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1 || (grid->np1 != 1 && grid->np1 != 2))
    return ASAM_E_INVALID;
  long N = (long)grid->nx * grid->ny * grid->nz;
  if (N > (1L << 31)) return ASAM_E_INVALID;
  {
    // every reduction launch must fit the partial-sum slots
    Geom tg;
    tg.nx = grid->nx; tg.ny = grid->ny; tg.nz = grid->nz; tg.np1 = grid->np1; tg.N = N;
    tg.fa = 1.0;
    if (apply_blocks(tg, 0) > kRedBlocks || apply_blocks(tg, 1) > kRedBlocks ||
        line_blocks(tg) > kRedBlocks ||
        (long)((grid->nx + 63) / 64) * ((grid->ny + 3) / 4) > kRedBlocks)
      return ASAM_E_INVALID;
  }
  asam_sound_ctx* c = new asam_sound_ctx();
  c->grid = *grid;
  asam_sound_default_control(&c->ctl);
  c->g.nx = grid->nx; c->g.ny = grid->ny; c->g.nz = grid->nz; c->g.np1 = grid->np1;
  c->g.N = N; c->g.fa = grid->fac_anela;
  c->n = N * grid->np1;
  int rc = 0;
  auto fin = [&](int code) { *out = c; return code; };
  if (hipGetDevice(&c->device) != hipSuccess) { c->err = "no HIP device"; return fin(ASAM_E_HIP); }
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(c, hipEventCreate(&c->ev0));
  HIPCHK(c, hipEventCreate(&c->ev1));
  const long P = (long)grid->np1 * grid->np1 * fac_cells(c->g);
  const long PL = (long)lower_words(grid->np1) * fac_cells(c->g);
  auto coef64 = [&](Coef<double>& k) -> int {
    int r = 0;
    r |= dalloc(c, &k.kx, N); r |= dalloc(c, &k.ky, N); r |= dalloc(c, &k.kz, N);
    if (grid->np1 == 2) {
      r |= dalloc(c, &k.tx, N); r |= dalloc(c, &k.ty, N); r |= dalloc(c, &k.tz, N);
      r |= dalloc(c, &k.gz, N);
    }
    r |= dalloc(c, &k.vol, N); r |= dalloc(c, &k.dut, N);
    r |= dalloc(c, &k.pl, PL); r |= dalloc(c, &k.pt, P); r |= dalloc(c, &k.pw, P);
    return r;
  };
  auto coef32 = [&](Coef<float>& k) -> int {
    int r = 0;
    r |= dalloc(c, &k.kx, N); r |= dalloc(c, &k.ky, N); r |= dalloc(c, &k.kz, N);
    if (grid->np1 == 2) {
      r |= dalloc(c, &k.tx, N); r |= dalloc(c, &k.ty, N); r |= dalloc(c, &k.tz, N);
      r |= dalloc(c, &k.gz, N);
      r |= dalloc(c, &k.pkz, 4 * N);
    }
    r |= dalloc(c, &k.vol, N); r |= dalloc(c, &k.dut, N);
    r |= dalloc(c, &k.pl, PL); r |= dalloc(c, &k.pt, P); r |= dalloc(c, &k.pw, P);
    return r;
  };
  rc |= coef64(c->c64);
  rc |= coef32(c->c32);
  long nf = (long)(grid->nx + 1) * grid->ny * grid->nz + (long)grid->nx * (grid->ny + 1) * grid->nz +
            (long)grid->nx * grid->ny * (grid->nz + 1);
  rc |= dalloc(c, &c->thfaces, nf);
  rc |= dalloc(c, &c->partial, 2 * kPartialRegion);
  rc |= dalloc(c, &c->tail_ctr, 64);
  rc |= dalloc(c, &c->st, 2);  // the state, and a copy the folded tails read (st + 1)
  rc |= dalloc(c, &c->dead, N);
  const long n = c->n;
  double** v64[] = {&c->r, &c->rt, &c->p, &c->ph, &c->v, &c->rh, &c->t, &c->xd, &c->bd, &c->xalt};
  for (double** q : v64) rc |= dalloc(c, q, n);
  float** v32[] = {&c->r32, &c->rt32, &c->p32, &c->ph32, &c->v32, &c->rh32, &c->t32, &c->d32};
  for (float** q : v32) rc |= dalloc(c, q, n);
  if (rc) return fin(ASAM_E_NOMEM);
  HIPCHK(c, hipHostMalloc((void**)&c->st_host, sizeof(KState), hipHostMallocDefault));
  memset(c->st_host, 0, sizeof(KState));
  HIPCHK(c, hipMemsetAsync(c->tail_ctr, 0, 64 * sizeof(unsigned int), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return fin(0);
}

extern "C" int asam_sound_destroy(asam_sound_ctx* c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->st_host) (void)hipHostFree(c->st_host);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

extern "C" int asam_sound_set_control(asam_sound_ctx* c, const asam_sound_control* ctl) {
  if (!c || !ctl) return ASAM_E_INVALID;
  if (ctl->method < 0 || ctl->method > 3) return fail(c, ASAM_E_INVALID, "method");
  if (ctl->step_gmres < 1 || ctl->step_gmres > kMaxM) return fail(c, ASAM_E_INVALID, "step_gmres");
  if (ctl->precision < 0 || ctl->precision > 1) return fail(c, ASAM_E_INVALID, "precision");
  if (ctl->precision == ASAM_PREC_MIXED && ctl->method == ASAM_METHOD_GMRES)
    return fail(c, ASAM_E_INVALID,
                "mixed precision runs a BiCGStab, FGMRES or FGMRESV inner solver");
  c->ctl = *ctl;
  return 0;
}

// ------------------------------------------------------------------ set-up
static int setup_common(asam_sound_ctx* c, const double* FU, const double* FV, const double* FW,
                        const double* V, const double* dx, const double* dy, const double* dz,
                        const double* rho, const double* rth, const double* T,
                        const double* rhov, const double* rhol, const double* rhor,
                        const double* DTU, const double* DTV, const double* DTW,
                        const double* DUT) {
  const Geom& g = c->g;
  const long N = g.N;
  const long nfu = (long)(g.nx + 1) * g.ny * g.nz, nfv = (long)g.nx * (g.ny + 1) * g.nz,
             nfw = (long)g.nx * g.ny * (g.nz + 1);
  for (long q = 0; q < N; ++q)
    if (!(V[q] >= 0.0) || !isfinite(V[q])) return fail(c, ASAM_E_NAN, "VolC must be finite, >= 0");
  double *dFU, *dFV, *dFW, *dV, *ddx, *ddy, *ddz, *tmp = nullptr;
  int rc = 0;
  rc |= dalloc(c, &dFU, nfu); rc |= dalloc(c, &dFV, nfv); rc |= dalloc(c, &dFW, nfw);
  rc |= dalloc(c, &dV, N); rc |= dalloc(c, &ddx, g.nx); rc |= dalloc(c, &ddy, g.ny);
  rc |= dalloc(c, &ddz, g.nz);
  if (rc) return ASAM_E_NOMEM;
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(dFU, FU, nfu * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dFV, FV, nfv * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dFW, FW, nfw * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dV, V, N * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(ddx, dx, g.nx * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(ddy, dy, g.ny * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(ddz, dz, g.nz * 8, hipMemcpyHostToDevice, s));
  double* thu = c->thfaces;
  double* thv = thu + nfu;
  double* thw = thv + nfv;
  if (rho) {
    double *drho, *drth, *dT, *drv = nullptr, *drl = nullptr, *drr = nullptr;
    rc |= dalloc(c, &drho, N); rc |= dalloc(c, &drth, N); rc |= dalloc(c, &dT, N);
    rc |= dalloc(c, &tmp, N);
    if (rhov) rc |= dalloc(c, &drv, N);
    if (rhol) rc |= dalloc(c, &drl, N);
    if (rhor) rc |= dalloc(c, &drr, N);
    if (rc) return ASAM_E_NOMEM;
    HIPCHK(c, hipMemcpyAsync(drho, rho, N * 8, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(drth, rth, N * 8, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(dT, T, N * 8, hipMemcpyHostToDevice, s));
    if (drv) HIPCHK(c, hipMemcpyAsync(drv, rhov, N * 8, hipMemcpyHostToDevice, s));
    if (drl) HIPCHK(c, hipMemcpyAsync(drl, rhol, N * 8, hipMemcpyHostToDevice, s));
    if (drr) HIPCHK(c, hipMemcpyAsync(drr, rhor, N * 8, hipMemcpyHostToDevice, s));
    launch_weights(s, g, drho, drth, dT, drv, drl, drr, tmp, c->c64.dut);
    launch_face_theta(s, g, dV, tmp, thu, thv, thw);
    HIPCHK(c, hipStreamSynchronize(s));
    dfree(c, drho); dfree(c, drth); dfree(c, dT); dfree(c, tmp);
    if (drv) dfree(c, drv);
    if (drl) dfree(c, drl);
    if (drr) dfree(c, drr);
  } else {
    HIPCHK(c, hipMemcpyAsync(thu, DTU, nfu * 8, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(thv, DTV, nfv * 8, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(thw, DTW, nfw * 8, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->c64.dut, DUT, N * 8, hipMemcpyHostToDevice, s));
  }
  launch_coef(s, g, c->grid.grad_full, c->grid.beta0_dt, c->grid.grav, dFU, dFV, dFW, dV, ddx,
              ddy, ddz, thu, thv, thw, c->c64);
  launch_factor(s, g, c->c64);
  launch_convert(s, g, c->c64, c->c32);
  launch_dead_mask(s, g, c->c64, c->dead);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(s));
  dfree(c, dV);
  // the face areas and spacings stay for the right-hand side and the post-solve update
  dfree(c, c->fu); dfree(c, c->fv); dfree(c, c->fw);
  dfree(c, c->dxv); dfree(c, c->dyv); dfree(c, c->dzv);
  c->fu = dFU; c->fv = dFV; c->fw = dFW; c->dxv = ddx; c->dyv = ddy; c->dzv = ddz;
  c->ready = 1;
  return 0;
}

extern "C" int asam_sound_setup(asam_sound_ctx* c, const double* FU, const double* FV,
                                const double* FW, const double* V, const double* dx,
                                const double* dy, const double* dz, const double* rho,
                                const double* rth, const double* T) {
  if (!c) return ASAM_E_INVALID;
  if (!FU || !FV || !FW || !V || !dx || !dy || !dz || !rho || !rth || !T)
    return fail(c, ASAM_E_INVALID, "null input");
  HIPCHK(c, hipSetDevice(c->device));
  return setup_common(c, FU, FV, FW, V, dx, dy, dz, rho, rth, T, nullptr, nullptr, nullptr,
                      nullptr, nullptr, nullptr, nullptr);
}

extern "C" int asam_sound_setup_moist(asam_sound_ctx* c, const double* FU, const double* FV,
                                      const double* FW, const double* V, const double* dx,
                                      const double* dy, const double* dz, const double* rho,
                                      const double* rth, const double* T, const double* rhov,
                                      const double* rhol, const double* rhor) {
  if (!c) return ASAM_E_INVALID;
  if (!FU || !FV || !FW || !V || !dx || !dy || !dz || !rho || !rth || !T)
    return fail(c, ASAM_E_INVALID, "null input");
  HIPCHK(c, hipSetDevice(c->device));
  return setup_common(c, FU, FV, FW, V, dx, dy, dz, rho, rth, T, rhov, rhol, rhor, nullptr,
                      nullptr, nullptr, nullptr);
}

extern "C" int asam_sound_setup_weights(asam_sound_ctx* c, const double* FU, const double* FV,
                                        const double* FW, const double* V, const double* dx,
                                        const double* dy, const double* dz, const double* DTU,
                                        const double* DTV, const double* DTW,
                                        const double* DUT) {
  if (!c) return ASAM_E_INVALID;
  if (!FU || !FV || !FW || !V || !dx || !dy || !dz || !DTU || !DTV || !DTW || !DUT)
    return fail(c, ASAM_E_INVALID, "null input");
  HIPCHK(c, hipSetDevice(c->device));
  return setup_common(c, FU, FV, FW, V, dx, dy, dz, nullptr, nullptr, nullptr, nullptr, nullptr,
                      nullptr, DTU, DTV, DTW, DUT);
}

// ------------------------------------------------------------------ enqueue helpers
namespace {

// Algorithmic bytes of one launch: every distinct array element read or written once.
double apply_bytes(const Geom& g, const ApplyReq& r) {
  const double N = (double)g.N, NP = g.np1;
  const double wc = r.prec ? 4 : 8, coefw = g.np1 == 1 ? 5 : 9;
  const double sx = r.vx ? 4 : 8, so = r.vo ? 4 : 8, su = r.vu ? 4 : 8;
  double b = N * coefw * wc + N * NP * (sx + so);
  if (r.dx) b += N * NP * 4;
  if (r.xw) b += N * NP * 8;
  if (r.b) b += N * NP * 8;
  if (r.y2) b += N * NP * so;
  if (r.dots & DOT_YU) b += N * NP * su;
  return b;
}
double line_bytes(const Geom& g, const LineReq& r, bool first) {
  const double N = (double)g.N, NP = g.np1;
  const double wc = r.prec ? 4 : 8, wv = r.vt ? 4 : 8;
  double vec = r.mode == LINE_PLAIN ? 2 : (r.mode == LINE_BICG_P ? (first ? 3 : 5) : 4);
  // factors: T^-1 and W (NP x NP words each) and L (lower_words) per cell
  return N * (2 * NP * NP + lower_words(g.np1)) * wc + N * NP * wv * vec;
}

struct Enq {
  asam_sound_ctx* c;
  hipStream_t s;
  int tr;
  int iter = 0;  // BiCGStab iteration index of the launches being enqueued (profiling)
  explicit Enq(asam_sound_ctx* cc, int t) : c(cc), s(cc->stream), tr(t) {}

  // The tails whose step touches only the state's scalars run inside the producing
  // kernel when the tuning asks for it; the others keep their finalize launch.
  bool fused(int kind) const {
    return c->g.tune.fused_tail == 1 &&
           (kind == FIN_BICG_INIT || kind == FIN_BICG_ALPHA || kind == FIN_BICG_HALF_OMEGA ||
            kind == FIN_BICG_END || kind == FIN_MIXED_OUTER);
  }
  KTail tail(int kind, int nq, int nblk, int nblk_hi) const {
    KTail t;
    t.kind = kind; t.nq = nq; t.nblk = nblk; t.nblk_hi = nblk_hi > 0 ? nblk_hi : nblk;
    t.partial = c->partial; t.st = c->st; t.counter = c->tail_ctr;
    return t;
  }
  // BiCGStab tails folded into the kernels that consume their scalars (tuning
  // fused_tail = 2): alpha into the s-step, the half-step test and omega into the update,
  // the exit test and beta into the next iteration's p-step.  The update then writes its
  // sums to the second partial region, and its tail stays pending until that p-step, or
  // until flush() launches it at the end of a batch of iterations.
  bool folding() const { return c->g.tune.fused_tail == 2; }
  bool end_pending = false;
  KTail folded(int kind, int nq, int nblk, int nblk_hi, const double* partial, int qmask) const {
    KTail t = tail(kind, nq, nblk, nblk_hi);
    t.partial = partial; t.qmask = qmask; t.counter = nullptr;
    return t;
  }
  void flush() {
    if (!end_pending) return;
    end_pending = false;
    fin(FIN_BICG_END, 2, blas_blocks(c->n), 0, 0, 0, 0, c->partial + kPartialRegion);
  }

  hipEvent_t ev() {
    if (c->ev_next >= c->ev_pool.size()) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_next++];
  }
  // Brackets one launch with events when profiling; tag 1: the launch is skipped by a
  // BiCGStab half-step exit, tag 2: it does the half-step x update (half_bytes).
  template <typename F>
  void timed(int cls, double bytes, F&& launch, int tag = 0, double half_bytes = 0.0) {
    if (!c->prof_on) { launch(); return; }
    hipEvent_t e0 = ev(), e1 = ev();
    if (!e0 || !e1) { launch(); return; }
    (void)hipEventRecord(e0, s);
    launch();
    (void)hipEventRecord(e1, s);
    c->prof_pending.push_back({cls, tag, iter, bytes, half_bytes, e0, e1});
  }

  // Launch a kernel and (kind >= 0) the tail of its reduction: fused into the kernel, or
  // a finalize launch after it.  fnblk_hi: workgroups behind partial slots 2-3.
  void apply(int prec, int vx, int vo, int vu, const void* x, void* y, const double* b,
             void* y2, const void* u, int dots, const int* done, const float* dx = nullptr,
             double* xw = nullptr, int tag = 0, int fkind = -1, int fnq = 0,
             int fnblk_hi = 0) {
    ApplyReq r;
    r.tr = tr; r.prec = prec; r.vx = vx; r.vo = vo; r.vu = vu; r.x = x; r.y = y; r.b = b;
    r.y2 = y2; r.u = u; r.dots = dots; r.partial = c->partial; r.done = done; r.dx = dx;
    r.xw = xw;
    const bool fuse = fkind >= 0 && dots && fused(fkind);
    if (fuse) r.tail = tail(fkind, fnq, apply_blocks(c->g, prec), fnblk_hi);
    timed(prec ? 1 : 0, apply_bytes(c->g, r),
          [&] { c->launch_err |= launch_apply(s, c->g, c->c64, c->c32, r) != 0; },
          tag);
    c->counts[0]++;
    if (fkind >= 0 && !fuse) fin(fkind, fnq, apply_blocks(c->g, prec), 0, 0, tag, fnblk_hi);
  }
  // Initial residual y = y2 = b - Op(x), y.y into the partial sums, then finalize fkind.
  // Unprofiled, a scan of x picks on the device which of two launches writes y: the full
  // FP64 apply, or, for an all-zero x (the usual first guess), the x = 0 form that reads
  // only b and the dead-cell mask (same values and partial sums, a quarter of the bytes).
  // Needs st->xnz = 0 and st->xz = 1 in the uploaded state.  dots: DOT_YY or 0; fkind < 0:
  // no finalize.
  void resid_init(int vo, const double* x, void* y, const double* b, void* y2, int dots,
                  int fkind) {
    if (c->prof_on) {
      apply(0, 0, vo, 0, x, y, b, y2, nullptr, dots, nullptr, nullptr, nullptr, 0, fkind, 1);
      return;
    }
    launch_any_nonzero(s, c->n, x, &c->st->xnz, &c->st->xz);
    ApplyReq r;
    r.tr = tr; r.prec = 0; r.vx = 0; r.vo = vo; r.vu = 0; r.x = x; r.y = y; r.b = b;
    r.y2 = y2; r.dots = dots; r.partial = c->partial; r.done = &c->st->xz;
    // exactly one of the two launches runs (the other returns on its flag), so at most
    // one of them arrives at the fused tail
    const bool fuse = fkind >= 0 && dots && fused(fkind);
    if (fuse) r.tail = tail(fkind, 1, apply_blocks(c->g, 0), 0);
    c->launch_err |= launch_apply(s, c->g, c->c64, c->c32, r) != 0;
    r.x = nullptr; r.x0_dead = c->dead; r.done = &c->st->xnz;
    c->launch_err |= launch_apply(s, c->g, c->c64, c->c32, r) != 0;
    c->counts[0]++;
    if (fkind >= 0 && !fuse) fin(fkind, 1, apply_blocks(c->g, 0));
  }
  void line(int prec, int mode, const void* src, void* p, const void* v, void* out,
            const double* scale, const int* done, int qslot = 0, const KTail& fold = KTail()) {
    LineReq r;
    r.tr = tr; r.prec = prec; r.vt = prec; r.mode = mode; r.src = src; r.p = p; r.v = v;
    r.out = out; r.scale = scale; r.st = c->st; r.partial = c->partial; r.done = done;
    r.qslot = qslot; r.fold = fold;
    timed(2 + prec, line_bytes(c->g, r, iter == 1),
          [&] { launch_line(s, c->g, c->c64, c->c32, r); });
    c->counts[1]++;
  }
  void fin(int kind, int nq, int nblk, int k = 0, int i = 0, int tag = 0, int nblk_hi = 0,
           const double* base = nullptr) {
    const double* p = base ? base : c->partial;
    timed(6, 8.0 * nq * nblk,
          [&] { launch_finalize(s, kind, nq, nblk, p, c->st, k, i, nullptr, nblk_hi); }, tag);
    c->counts[3]++;
  }
  void bicg_update(int vt, void* x, const void* ph, const void* rh, void* r, const void* t,
                   const void* rt) {
    const double w = vt ? 4 : 8, nv = (double)c->n;
    const bool fuse = fused(FIN_BICG_END);
    const KTail tl = fuse ? tail(FIN_BICG_END, 2, blas_blocks(c->n), 0) : KTail();
    const bool fold = folding();
    // folded: the half-step test and omega of the preceding apply (slots 0-1) and s-step
    // (slot 2) are run by every workgroup; the sums go to the second region
    KTail fo;
    if (fold) {
      fo = folded(FIN_BICG_HALF_OMEGA, 3, apply_blocks(c->g, vt), line_blocks(c->g), c->partial,
                  0x7);
      fo.copy = c->st + 1;  // the exit test folded into the next p-step reads this copy
    }
    double* part = fold ? c->partial + kPartialRegion : c->partial;
    timed(4 + vt, 8 * nv * w,
          [&] { launch_bicg_update(s, c->n, vt, x, ph, rh, r, t, rt, c->st, part, tl, fo); },
          2, 3 * nv * w);
    c->counts[2]++;
    if (fold) end_pending = true;
    else if (!fuse) fin(FIN_BICG_END, 2, blas_blocks(c->n));
  }
  // classical Gram-Schmidt: nq projections of w on basis vectors k0.., then the removal
  // of all i + 1 projections with |w|^2
  void cgs_dots(int vt, void* w, void* const* vs, int k0, int nq, int i) {
    const double nv = (double)c->n, ws = vt ? 4 : 8;
    timed(4 + vt, (nq + 1) * nv * ws,
          [&] { launch_cgs_dots(s, c->n, vt, w, vs, k0, nq, c->st, c->partial); });
    c->counts[2]++;
    fin(FIN_GM_HV, nq, blas_blocks(c->n), k0, i);
  }
  void cgs_update(int vt, void* w, void* const* vs, int i) {
    const double nv = (double)c->n, ws = vt ? 4 : 8;
    timed(4 + vt, (i + 3) * nv * ws,
          [&] { launch_cgs_update(s, c->n, vt, w, vs, i, c->st, c->partial); });
    c->counts[2]++;
    fin(FIN_GM_LAST, 1, blas_blocks(c->n), 0, i);
  }
  void mgs(int vt, void* w, const void* vprev, const void* vk, int k, int i, int chk = 0) {
    const double ws = vt ? 4 : 8, nv = (double)c->n;
    double vecs = (k > 0 ? 3 : 1) + (k == i + 1 ? 0 : 1);
    timed(4 + vt, vecs * nv * ws,
          [&] { launch_mgs(s, c->n, vt, w, vprev, vk, k, i, c->st, c->partial, chk); });
    c->counts[2]++;
  }
  // Steps k = 1..i+1 of the modified Gram-Schmidt chain of Arnoldi step i (w = vs[i+1]),
  // H(0,i) already in the state.  Step k >= 2 sums the previous step's partial products
  // itself (the same fixed order in every workgroup), so only the last step ends in a
  // finalize; the steps alternate between partial slots 2 and 3.
  void mgs_chain(int vt, void* const* vs, int i) {
    const double ws = vt ? 4 : 8, nv = (double)c->n;
    const double* prev = nullptr;
    int slot = 2;
    for (int k = 1; k <= i + 1; ++k) {
      const bool self = k == i + 1;
      double* out = c->partial + (long)slot * kRedBlocks;
      timed(4 + vt, (self ? 3 : 4) * nv * ws, [&] {
        launch_mgs(s, c->n, vt, vs[i + 1], vs[k - 1], self ? nullptr : vs[k], k, i, c->st, out,
                   0, prev);
      });
      c->counts[2]++;
      prev = out;
      slot ^= 1;
    }
    fin(FIN_GM_LAST, 1, blas_blocks(c->n), 0, i, 0, 0, prev);
  }
};

// Resolve recorded launches into the per-class sums.  last_iter > 0: launches of
// BiCGStab iterations after it returned at once (the solve had converged) and are
// dropped; half: iteration last_iter ended with a half-step exit.
void prof_resolve(asam_sound_ctx* c, bool half, int last_iter = 0) {
  for (auto& r : c->prof_pending) {
    if (last_iter > 0 && r.iter > last_iter) continue;
    const bool in_last = last_iter > 0 && r.iter == last_iter;
    if (half && in_last && r.tag == 1) continue;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.e0, r.e1) != hipSuccess) continue;
    c->prof_ms[r.cls] += ms;
    c->prof_bytes[r.cls] += (half && in_last && r.tag == 2) ? r.half_bytes : r.bytes;
    c->prof_launches[r.cls] += 1;
  }
  c->prof_pending.clear();
  c->ev_next = 0;
}

int sync_state(asam_sound_ctx* c) {
  HIPCHK(c, hipMemcpyAsync(c->st_host, c->st, sizeof(KState), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int upload_state(asam_sound_ctx* c) {
  HIPCHK(c, hipMemcpyAsync(c->st, c->st_host, sizeof(KState), hipMemcpyHostToDevice, c->stream));
  return 0;
}

// One BiCGStab iteration (SpMatrix4CellPar_Mod.f90:280-337) as 5 fused kernels; the
// scalar tails (alpha; half-step test and omega; exit tests and beta) run as finalize
// launches (fused_tail = 0), in the last workgroup of the kernel before them (1), or in
// every workgroup of the kernel after them (2; the exit tests and beta of the last
// iteration of a batch then run at the batch's flush()).
// prec/vt: 0 FP64, 1 FP32 vectors and coefficients (dots always FP64).
template <typename VT>
void bicg_iteration(Enq& e, int prec, VT* x, VT* r, VT* rt, VT* p, VT* ph, VT* v, VT* rh, VT* t) {
  asam_sound_ctx* c = e.c;
  KState* st = c->st;
  const Geom& g = c->g;
  const int vt = prec;
  const bool fold = e.folding();
  // p = r + beta (p - omega v) (p = r at iter 1); pHat = PSolve(p).  Folded: the previous
  // iteration's exit test and beta first (the update's sums, second region)
  KTail fend;
  if (fold && e.end_pending) {
    fend = e.folded(FIN_BICG_END, 2, blas_blocks(c->n), 0, c->partial + kPartialRegion, 0x3);
    fend.src = c->st + 1;  // the update's copy: the step rewrites rho, iter and done
    e.end_pending = false;
  }
  e.line(prec, LINE_BICG_P, r, p, v, ph, nullptr, &st->done, 0, fend);
  // v = MatVec(pHat); alpha = rho / (rTilde.v) (folded: in the s-step)
  e.apply(prec, vt, vt, vt, ph, v, nullptr, nullptr, rt, DOT_YU, &st->done, nullptr, nullptr, 0,
          fold ? -1 : FIN_BICG_ALPHA, 2);
  // s = r - alpha v (into r); |s| (partial slot 2); rHat = PSolve(s)
  e.line(prec, LINE_BICG_S, nullptr, r, v, rh, nullptr, &st->done, 2,
         fold ? e.folded(FIN_BICG_ALPHA, 2, apply_blocks(g, prec), 0, c->partial, 0x2)
              : KTail());
  // t = MatVec(rHat); t.t, t.s; then the half-step exit test on |s| or omega = (t.s)/(t.t)
  // (folded: in the update)
  e.apply(prec, vt, vt, vt, rh, t, nullptr, nullptr, r, DOT_YY | DOT_YU, &st->done, nullptr,
          nullptr, 0, fold ? -1 : FIN_BICG_HALF_OMEGA, 3, line_blocks(g));
  // x += alpha pHat + omega rHat; r = s - omega t (half exit: x += alpha pHat);
  // |r|, rTilde.r; exit tests, beta
  e.bicg_update(vt, x, ph, rh, r, t, rt);
}

// Runs BiCGStab iterations until the device reports done; returns 0 or error.
template <typename VT>
int bicg_loop(Enq& e, int prec, int maxiter, VT* x, VT* r, VT* rt, VT* p, VT* ph, VT* v, VT* rh,
              VT* t, int* iters_out, int first_chunk) {
  asam_sound_ctx* c = e.c;
  int issued = 0;
  // First chunk: the iteration count the same solve needed last time; then one at a time.
  int chunk = first_chunk > 0 ? first_chunk : 4;
  if (chunk > maxiter) chunk = maxiter;
  while (issued < maxiter) {
    int q = chunk;
    if (issued + q > maxiter) q = maxiter - issued;
    for (int k = 0; k < q; ++k) {
      e.iter = issued + k + 1;
      bicg_iteration<VT>(e, prec, x, r, rt, p, ph, v, rh, t);
    }
    issued += q;
    e.flush();  // the last iteration's exit test, when folded tails left it pending
    int rc = sync_state(c);
    if (rc) return rc;
    if (c->prof_on) {
      const bool dn = c->st_host->done != 0;
      prof_resolve(c, dn && c->st_host->half != 0, dn ? c->st_host->iter : 0);
    }
    if (c->st_host->done) break;
    chunk = 1;
  }
  e.iter = 0;  // launches after the loop belong to no BiCGStab iteration
  *iters_out = issued;
  return 0;
}

}  // namespace

static int solve_bicg_fp64(asam_sound_ctx* c, int tr, double* x, const double* b, int* maxiter,
                           double* tol, asam_sound_info* info) {
  Enq e(c, tr);
  KState* hs = c->st_host;
  memset(hs, 0, sizeof(KState));
  hs->maxiter = *maxiter;
  hs->tol = *tol;
  hs->xz = 1;
  int rc = upload_state(c);
  if (rc) return rc;
  // r = b - MatVec(x); rTilde = r; rNorm0 = |r|
  e.resid_init(0, x, c->r, b, c->rt, DOT_YY, FIN_BICG_INIT);
  int issued = 0;
  if (*maxiter > 0) {
    e.iter = 0;
    rc = bicg_loop<double>(e, 0, *maxiter, x, c->r, c->rt, c->p, c->ph, c->v, c->rh, c->t,
                           &issued, c->last_iters);
    if (rc) return rc;
  } else {
    rc = sync_state(c);
    if (rc) return rc;
  }
  const KState& s = *hs;
  int res = s.done ? s.result : 0;
  int it = res < 0 ? -res : (s.status == 1 ? res / 2 : res);
  if (s.status == 2) it = res;
  c->last_iters = it > 0 ? it : 1;
  *maxiter = res;
  *tol = s.rnorm;
  if (info) {
    info->iterations = it;
    info->restarts = 0;
    info->status = s.status == 0 ? ASAM_CONVERGED : (s.status == 2 ? ASAM_BREAKDOWN : ASAM_NOT_CONVERGED);
    info->abs_residual = s.rnorm;
    info->rel_residual = s.rnorm0 > 0 ? s.rnorm / s.rnorm0 : 0.0;
  }
  return 0;
}

static int ensure_gmres(asam_sound_ctx* c, int m) {
  if (c->gm_m >= m) return 0;
  for (double* q : c->gv) dfree(c, q);
  for (double* q : c->gz) dfree(c, q);
  c->gv.assign(m + 1, nullptr);
  c->gz.assign(m, nullptr);
  int rc = 0;
  for (int i = 0; i <= m; ++i) rc |= dalloc(c, &c->gv[i], c->n);
  for (int i = 0; i < m; ++i) rc |= dalloc(c, &c->gz[i], c->n);
  if (!c->gtmp) rc |= dalloc(c, &c->gtmp, c->n);
  if (rc) return ASAM_E_NOMEM;
  c->gm_m = m;
  return 0;
}

static int ensure_gmres32(asam_sound_ctx* c, int m) {
  if (c->gm32_m >= m) return 0;
  for (float* q : c->gv32) dfree(c, q);
  for (float* q : c->gz32) dfree(c, q);
  c->gv32.assign(m + 1, nullptr);
  c->gz32.assign(m, nullptr);
  int rc = 0;
  for (int i = 0; i <= m; ++i) rc |= dalloc(c, &c->gv32[i], c->n);
  for (int i = 0; i < m; ++i) rc |= dalloc(c, &c->gz32[i], c->n);
  if (rc) return ASAM_E_NOMEM;
  c->gm32_m = m;
  return 0;
}

// FGMRES(m) (Iter_Mod.f90:121-199), FGMRESV(m) (:201-288: FGMRES with classical
// Gram-Schmidt) or left-preconditioned GMRES(m) (:32-119), FP64.
static int solve_gmres_fp64(asam_sound_ctx* c, int tr, int flexible, int classical, double* x,
                            const double* b, int maxcycles, double tol,
                            asam_sound_info* info) {
  const int m = c->ctl.step_gmres;
  int rc = ensure_gmres(c, m);
  if (rc) return rc;
  Enq e(c, tr);
  KState* hs = c->st_host;
  memset(hs, 0, sizeof(KState));
  hs->m = m;
  hs->gtol = tol;
  hs->maxcycles = maxcycles;
  hs->xz = 1;
  rc = upload_state(c);
  if (rc) return rc;
  KState* st = c->st;
  const Geom& g = c->g;
  const long n = c->n;
  for (int cyc = 0; cyc < maxcycles; ++cyc) {
    // the first cycle's residual may start from a zero guess; later ones from the iterate
    if (flexible) {
      if (cyc == 0)
        e.resid_init(0, x, c->gv[0], b, nullptr, DOT_YY, -1);
      else
        e.apply(0, 0, 0, 0, x, c->gv[0], b, nullptr, nullptr, DOT_YY, &st->done);
      e.fin(FIN_GM_INIT, 1, apply_blocks(g, 0));
    } else {
      if (cyc == 0)
        e.resid_init(0, x, c->gtmp, b, nullptr, 0, -1);
      else
        e.apply(0, 0, 0, 0, x, c->gtmp, b, nullptr, nullptr, 0, &st->done);
      e.line(0, LINE_PLAIN, c->gtmp, nullptr, nullptr, c->gv[0], nullptr, &st->done);
      e.mgs(0, c->gv[0], nullptr, nullptr, 0, -1, 1);
      e.fin(FIN_GM_INIT, 1, blas_blocks(n));
    }
    for (int i = 0; i < m; ++i) {
      if (flexible) {
        e.line(0, LINE_PLAIN, c->gv[i], nullptr, nullptr, c->gz[i], &st->vscale[i],
               &st->cycle_done);
        e.apply(0, 0, 0, 0, c->gz[i], c->gv[i + 1], nullptr, nullptr, c->gv[0], DOT_YU,
                &st->cycle_done);
        e.fin(FIN_GM_H, 2, apply_blocks(g, 0), 0, i);
      } else {
        e.apply(0, 0, 0, 0, c->gv[i], c->gtmp, nullptr, nullptr, nullptr, 0, &st->cycle_done);
        e.line(0, LINE_PLAIN, c->gtmp, nullptr, nullptr, c->gv[i + 1], &st->vscale[i],
               &st->cycle_done);
        e.mgs(0, c->gv[i + 1], nullptr, c->gv[0], 0, i);
        e.fin(FIN_GM_H, 1, blas_blocks(n), 0, i);
      }
      if (classical) {
        // H(0,i) came with the apply; the other projections use the same w
        void* const* vs = (void* const*)c->gv.data();
        for (int k0 = 1; k0 <= i; k0 += 4)
          e.cgs_dots(0, c->gv[i + 1], vs, k0, i + 1 - k0 < 4 ? i + 1 - k0 : 4, i);
        e.cgs_update(0, c->gv[i + 1], vs, i);
        continue;
      }
      e.mgs_chain(0, (void* const*)c->gv.data(), i);
    }
    launch_gm_backsolve(e.s, st);
    if (flexible)
      launch_gmres_update(e.s, n, 0, x, (void* const*)c->gz.data(), m, 0, st);
    else
      launch_gmres_update(e.s, n, 0, x, (void* const*)c->gv.data(), m, 1, st);
    c->counts[2]++;
    rc = sync_state(c);
    if (rc) return rc;
    if (hs->done) break;
  }
  if (info) {
    info->iterations = hs->total;
    info->restarts = hs->cycles;
    info->status = (hs->grho <= tol * hs->grho0) ? ASAM_CONVERGED : ASAM_NOT_CONVERGED;
    info->abs_residual = hs->grho;
    info->rel_residual = hs->grho0 > 0 ? hs->grho / hs->grho0 : 0.0;
  }
  return 0;
}

// FP32-inner BiCGStab with FP64 iterative refinement.
//
// Refinement step k is the FP64 residual of x_k (with x_k = x_{k-1} + d for k > 0) and,
// unless it meets the tolerance, an FP32 inner solve of A d = r.  x_k lives in
// bufs[k & 1]; the device records the parity of the last step (xpar).  When the previous
// solve of this context left its per-step inner iteration counts, the whole schedule of
// that many steps is enqueued at once and synchronised once: kernels after convergence
// return at once on the device flags, and an inner solve the schedule cuts short only
// makes the next refinement step start from a less refined x.  Whatever is not finished
// then continues step by step.  The caller's maxiter (the BiCGStab iteration cap of the
// FP64 solve) caps every inner solve, together with the control's max_inner.
static int solve_mixed(asam_sound_ctx* c, int tr, double* x, const double* b, double tol,
                       int maxiter, asam_sound_info* info) {
  Enq e(c, tr);
  KState* hs = c->st_host;
  memset(hs, 0, sizeof(KState));
  const int max_inner = maxiter < c->ctl.max_inner ? maxiter : c->ctl.max_inner;
  const int max_outer = c->ctl.max_outer;
  hs->outer_tol = tol;
  hs->inner_tol = c->ctl.inner_tol;
  hs->maxiter = max_inner;
  hs->xz = 1;
  int rc = upload_state(c);
  if (rc) return rc;
  KState* st = c->st;
  double* const bufs[2] = {x, (x == c->xalt) ? c->xd : c->xalt};
  auto outer_step = [&](int k) {
    if (k == 0)
      e.resid_init(1, bufs[0], c->r32, b, c->rt32, DOT_YY, FIN_MIXED_OUTER);
    else
      e.apply(0, 0, 1, 0, bufs[(k - 1) & 1], c->r32, b, c->rt32, nullptr, DOT_YY,
              &st->outer_done, c->d32, bufs[k & 1], 0, FIN_MIXED_OUTER, 1);
  };
  auto inner_fill = [&] {
    e.timed(5, 4.0 * c->n, [&] { launch_fill(e.s, c->n, 1, c->d32, 0.0); });
  };
  int k = 0;  // last refinement step enqueued
  const int ns = c->mixed_steps;
  if (!c->prof_on && ns > 0 && ns <= 16 && ns <= max_outer && c->mixed_cap == max_inner) {
    outer_step(0);
    for (k = 1; k <= ns; ++k) {
      inner_fill();
      const int h = c->inner_hist[k - 1] > 0 ? c->inner_hist[k - 1] : 1;
      for (int q = 0; q < h && q < max_inner; ++q) {
        e.iter = q + 1;
        bicg_iteration<float>(e, 1, c->d32, c->r32, c->rt32, c->p32, c->ph32, c->v32, c->rh32,
                              c->t32);
      }
      e.iter = 0;
      e.flush();
      outer_step(k);
    }
    k = ns;
  } else {
    outer_step(0);
  }
  rc = sync_state(c);
  if (rc) return rc;
  bool converged = hs->outer_done != 0;
  // step-by-step continuation: k is the device's current step (its residual is known)
  k = hs->outer_done ? hs->outer : hs->outer - 1;
  while (!converged && k < max_outer) {
    inner_fill();
    int issued = 0;
    const int hist = k < 32 ? c->inner_hist[k] : 0;
    rc = bicg_loop<float>(e, 1, max_inner, c->d32, c->r32, c->rt32, c->p32, c->ph32,
                          c->v32, c->rh32, c->t32, &issued, hist > 0 ? hist : 4);
    if (rc) return rc;
    ++k;
    outer_step(k);
    rc = sync_state(c);
    if (rc) return rc;
    converged = hs->outer_done != 0;
  }
  // per-step inner counts as the device recorded them, for the next solve's schedule
  const int steps = hs->outer;
  for (int q = 0; q < steps && q < 16; ++q) c->inner_hist[q] = hs->hist[q] > 0 ? hs->hist[q] : 1;
  // a solve that ended on a non-finite residual leaves no schedule
  c->mixed_steps = (converged && isfinite(hs->outer_r)) ? steps : 0;
  c->mixed_cap = max_inner;
  double* xfin = bufs[hs->xpar & 1];
  if (xfin != x) HIPCHK(c, hipMemcpyAsync(x, xfin, c->n * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (info) {
    info->iterations = hs->inner_total;
    // without convergence the last residual's tail also set up a step that did not run
    info->restarts = converged ? steps : (steps > 0 ? steps - 1 : 0);
    info->status = converged ? ASAM_CONVERGED : ASAM_NOT_CONVERGED;
    info->abs_residual = hs->outer_r;
    info->rel_residual = hs->outer_r0 > 0 ? hs->outer_r / hs->outer_r0 : 0.0;
  }
  return 0;
}

// FP32-inner FGMRES / FGMRESV with FP64 iterative refinement.  Each refinement step is
// the FP64 residual r = b - Op(x) (FP64 coefficients; written as the FP32 first basis
// vector) and one FP32 cycle of at most m Arnoldi steps on A d = r from d = 0, to the inner
// tolerance; then x += d is fused into the next step's residual.  The reference's restart
// (a new cycle from the current iterate, Iter_Mod.f90:136-139) is the next refinement step,
// so every cycle starts from an FP64 residual.  The caller's maxiter (restart cycles of
// the FP64 solve) therefore caps the refinement steps, together with the control's
// max_outer.
static int solve_mixed_gm(asam_sound_ctx* c, int tr, int classical, double* x, const double* b,
                          double tol, int maxiter, asam_sound_info* info) {
  const int max_outer = maxiter < c->ctl.max_outer ? maxiter : c->ctl.max_outer;
  const int m = c->ctl.step_gmres;
  int rc = ensure_gmres32(c, m);
  if (rc) return rc;
  Enq e(c, tr);
  KState* hs = c->st_host;
  memset(hs, 0, sizeof(KState));
  hs->outer_tol = tol;
  hs->inner_tol = c->ctl.inner_tol;
  hs->m = m;
  hs->maxcycles = 1;
  hs->xz = 1;
  rc = upload_state(c);
  if (rc) return rc;
  KState* st = c->st;
  const Geom& g = c->g;
  const long n = c->n;
  double* const bufs[2] = {x, (x == c->xalt) ? c->xd : c->xalt};
  float* const v0 = c->gv32[0];
  auto outer_step = [&](int k) {
    if (k == 0)
      e.resid_init(1, bufs[0], v0, b, c->rt32, DOT_YY, FIN_MIXED_GM);
    else
      e.apply(0, 0, 1, 0, bufs[(k - 1) & 1], v0, b, c->rt32, nullptr, DOT_YY, &st->outer_done,
              c->d32, bufs[k & 1], 0, FIN_MIXED_GM, 1);
  };
  outer_step(0);
  rc = sync_state(c);
  if (rc) return rc;
  int k = 0;
  while (!hs->outer_done && k < max_outer) {
    e.timed(5, 4.0 * n, [&] { launch_fill(e.s, n, 1, c->d32, 0.0); });
    for (int i = 0; i < m; ++i) {
      e.line(1, LINE_PLAIN, c->gv32[i], nullptr, nullptr, c->gz32[i], &st->vscale[i],
             &st->cycle_done);
      e.apply(1, 1, 1, 1, c->gz32[i], c->gv32[i + 1], nullptr, nullptr, v0, DOT_YU,
              &st->cycle_done);
      e.fin(FIN_GM_H, 2, apply_blocks(g, 1), 0, i);
      if (classical) {
        void* const* vs = (void* const*)c->gv32.data();
        for (int k0 = 1; k0 <= i; k0 += 4)
          e.cgs_dots(1, c->gv32[i + 1], vs, k0, i + 1 - k0 < 4 ? i + 1 - k0 : 4, i);
        e.cgs_update(1, c->gv32[i + 1], vs, i);
        continue;
      }
      e.mgs_chain(1, (void* const*)c->gv32.data(), i);
    }
    launch_gm_backsolve(e.s, st);
    launch_gmres_update(e.s, n, 1, c->d32, (void* const*)c->gz32.data(), m, 0, st);
    c->counts[2]++;
    ++k;
    outer_step(k);
    rc = sync_state(c);
    if (rc) return rc;
  }
  const bool converged = hs->outer_done != 0;
  double* xfin = bufs[hs->xpar & 1];
  if (xfin != x) HIPCHK(c, hipMemcpyAsync(x, xfin, n * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (info) {
    info->iterations = hs->inner_total;
    // without convergence the last residual's tail also set up a step that did not run
    info->restarts = converged ? hs->outer : (hs->outer > 0 ? hs->outer - 1 : 0);
    info->status = converged ? ASAM_CONVERGED : ASAM_NOT_CONVERGED;
    info->abs_residual = hs->outer_r;
    info->rel_residual = hs->outer_r0 > 0 ? hs->outer_r / hs->outer_r0 : 0.0;
  }
  return 0;
}

static int solve_device_impl(asam_sound_ctx* c, int tr, double* x, const double* b, int* maxiter,
                             double* tol, asam_sound_info* info) {
  if (!c) return ASAM_E_INVALID;
  if (!c->ready) return fail(c, ASAM_E_STATE, "solve before setup");
  if (!x || !b || !maxiter || !tol) return fail(c, ASAM_E_INVALID, "null argument");
  if (!(*tol >= 0.0)) return fail(c, ASAM_E_INVALID, "tol");
  HIPCHK(c, hipSetDevice(c->device));
  memset(c->counts, 0, sizeof(c->counts));
  c->launch_err = 0;
  asam_sound_info tmp;
  memset(&tmp, 0, sizeof tmp);
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  int rc;
  if (c->ctl.precision == ASAM_PREC_MIXED && c->ctl.method != ASAM_METHOD_BICGSTAB) {
    rc = solve_mixed_gm(c, tr, c->ctl.method == ASAM_METHOD_FGMRESV, x, b, *tol, *maxiter,
                        &tmp);
  } else if (c->ctl.precision == ASAM_PREC_MIXED) {
    rc = solve_mixed(c, tr, x, b, *tol, *maxiter, &tmp);
  } else if (c->ctl.method == ASAM_METHOD_BICGSTAB) {
    rc = solve_bicg_fp64(c, tr, x, b, maxiter, tol, &tmp);
  } else {
    rc = solve_gmres_fp64(c, tr, c->ctl.method != ASAM_METHOD_GMRES,
                          c->ctl.method == ASAM_METHOD_FGMRESV, x, b, *maxiter, *tol,
                          &tmp);
  }
  if (rc) return rc;
  if (c->launch_err) {
    c->launch_err = 0;
    return fail(c, ASAM_E_INVALID, "an operator apply has no kernel for its option set");
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  if (c->prof_on) prof_resolve(c, false);
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  tmp.ms = ms;
  HIPCHK(c, hipGetLastError());
  if (!isfinite(tmp.abs_residual)) {
    if (info) *info = tmp;
    return fail(c, ASAM_E_NAN, "non-finite residual");
  }
  if (info) *info = tmp;
  return 0;
}

extern "C" int asam_sound_solve_device(asam_sound_ctx* c, double* x, const double* b,
                                       int* maxiter, double* tol, asam_sound_info* info) {
  return solve_device_impl(c, 0, x, b, maxiter, tol, info);
}

extern "C" int asam_sound_solve_T_device(asam_sound_ctx* c, double* x, const double* b,
                                         int* maxiter, double* tol, asam_sound_info* info) {
  return solve_device_impl(c, 1, x, b, maxiter, tol, info);
}

static int host_solve(asam_sound_ctx* c, int tr, double* x, const double* b, int* maxiter,
                      double* tol, asam_sound_info* info) {
  if (!c) return ASAM_E_INVALID;
  if (!x || !b) return fail(c, ASAM_E_INVALID, "null argument");
  if (!c->ready) return fail(c, ASAM_E_STATE, "solve before setup");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(c->xd, x, c->n * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->bd, b, c->n * 8, hipMemcpyHostToDevice, c->stream));
  int rc = solve_device_impl(c, tr, c->xd, c->bd, maxiter, tol, info);
  if (rc && rc != ASAM_E_NAN) return rc;
  HIPCHK(c, hipMemcpyAsync(x, c->xd, c->n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return rc;
}

extern "C" int asam_sound_solve(asam_sound_ctx* c, double* x, const double* b, int* maxiter,
                                double* tol, asam_sound_info* info) {
  return host_solve(c, 0, x, b, maxiter, tol, info);
}

extern "C" int asam_sound_solve_T(asam_sound_ctx* c, double* x, const double* b, int* maxiter,
                                  double* tol, asam_sound_info* info) {
  return host_solve(c, 1, x, b, maxiter, tol, info);
}

// ------------------------------------------------------------------ operator pieces
extern "C" int asam_sound_apply_device(asam_sound_ctx* c, double* y, const double* x, int tr) {
  if (!c) return ASAM_E_INVALID;
  if (!c->ready) return fail(c, ASAM_E_STATE, "apply before setup");
  if (!x || !y || x == y) return fail(c, ASAM_E_INVALID, "x, y must be distinct buffers");
  HIPCHK(c, hipSetDevice(c->device));
  c->launch_err = 0;
  Enq e(c, tr ? 1 : 0);
  e.apply(0, 0, 0, 0, x, y, nullptr, nullptr, nullptr, 0, nullptr);
  if (c->launch_err) {
    c->launch_err = 0;
    return fail(c, ASAM_E_INVALID, "an operator apply has no kernel for its option set");
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int asam_sound_psolve_device(asam_sound_ctx* c, double* y, const double* x, int tr) {
  if (!c) return ASAM_E_INVALID;
  if (!c->ready) return fail(c, ASAM_E_STATE, "psolve before setup");
  if (!x || !y) return fail(c, ASAM_E_INVALID, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  Enq e(c, tr ? 1 : 0);
  e.line(0, LINE_PLAIN, x, nullptr, nullptr, y, nullptr, nullptr);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

static int host_unary(asam_sound_ctx* c, double* y, const double* x, int tr, int which) {
  if (!c) return ASAM_E_INVALID;
  if (!x || !y) return fail(c, ASAM_E_INVALID, "null argument");
  if (!c->ready) return fail(c, ASAM_E_STATE, "before setup");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(c->xd, x, c->n * 8, hipMemcpyHostToDevice, c->stream));
  int rc = which == 0 ? asam_sound_apply_device(c, c->bd, c->xd, tr)
                      : asam_sound_psolve_device(c, c->bd, c->xd, tr);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(y, c->bd, c->n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int asam_sound_apply(asam_sound_ctx* c, double* y, const double* x) {
  return host_unary(c, y, x, 0, 0);
}
extern "C" int asam_sound_apply_T(asam_sound_ctx* c, double* y, const double* x) {
  return host_unary(c, y, x, 1, 0);
}
extern "C" int asam_sound_psolve(asam_sound_ctx* c, double* y, const double* x) {
  return host_unary(c, y, x, 0, 1);
}
extern "C" int asam_sound_psolve_T(asam_sound_ctx* c, double* y, const double* x) {
  return host_unary(c, y, x, 1, 1);
}

extern "C" int asam_sound_theta_vjp_device(asam_sound_ctx* c, double* grad, const double* mu,
                                           const double* x) {
  if (!c) return ASAM_E_INVALID;
  if (!c->ready) return fail(c, ASAM_E_STATE, "before setup");
  if (!grad || !mu || !x) return fail(c, ASAM_E_INVALID, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  launch_theta_vjp(c->stream, c->g, c->c64, c->thfaces, grad, mu, x);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// ------------------------------------------------------------------ device helpers
// ------------------------------------------------------------------ either side of the solve
static ProjInputs proj_inputs(const asam_sound_ctx* c) {
  ProjInputs in;
  const Geom& g = c->g;
  const long nfu = (long)(g.nx + 1) * g.ny * g.nz, nfv = (long)g.nx * (g.ny + 1) * g.nz;
  in.grad_full = c->grid.grad_full;
  in.beta0_dt = c->grid.beta0_dt;
  in.grav = c->grid.grav;
  in.FU = c->fu; in.FV = c->fv; in.FW = c->fw;
  in.thu = c->thfaces; in.thv = c->thfaces + nfu; in.thw = c->thfaces + nfu + nfv;
  in.V = c->c64.vol; in.dut = c->c64.dut;
  in.dx = c->dxv; in.dy = c->dyv; in.dz = c->dzv;
  return in;
}

extern "C" int asam_sound_rhs_device(asam_sound_ctx* c, double* b, const double* U,
                                     const double* V, const double* W, const double* incr_rt,
                                     const double* incr_r) {
  if (!c) return ASAM_E_INVALID;
  if (!c->ready) return fail(c, ASAM_E_STATE, "rhs before setup");
  if (!b || !U || !V || !W || !incr_rt || (c->g.np1 == 2 && !incr_r))
    return fail(c, ASAM_E_INVALID, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  launch_rhs(c->stream, c->g, proj_inputs(c), b, U, V, W, incr_rt, incr_r);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int asam_sound_update_device(asam_sound_ctx* c, double* x, double* U, double* V,
                                        double* W, double* incr_rt, double* incr_r) {
  if (!c) return ASAM_E_INVALID;
  if (!c->ready) return fail(c, ASAM_E_STATE, "update before setup");
  if (!x || !U || !V || !W || !incr_rt) return fail(c, ASAM_E_INVALID, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  launch_update(c->stream, c->g, proj_inputs(c), x, U, V, W, incr_rt, incr_r);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// Host-array forms: staged through one device block [x|b, U, V, W, incr_rt, incr_r].
static int pj_stage(asam_sound_ctx* c, double** xb, double** U, double** V, double** W,
                    double** it, double** ir) {
  const Geom& g = c->g;
  const long nfu = (long)(g.nx + 1) * g.ny * g.nz, nfv = (long)g.nx * (g.ny + 1) * g.nz,
             nfw = (long)g.nx * g.ny * (g.nz + 1);
  if (!c->pj_scratch) {
    int rc = dalloc(c, &c->pj_scratch, c->n + nfu + nfv + nfw + 2 * g.N);
    if (rc) return rc;
  }
  *xb = c->pj_scratch;
  *U = *xb + c->n;
  *V = *U + nfu;
  *W = *V + nfv;
  *it = *W + nfw;
  *ir = *it + g.N;
  return 0;
}

extern "C" int asam_sound_rhs(asam_sound_ctx* c, double* b, const double* U, const double* V,
                              const double* W, const double* incr_rt, const double* incr_r) {
  if (!c) return ASAM_E_INVALID;
  if (!c->ready) return fail(c, ASAM_E_STATE, "rhs before setup");
  if (!b || !U || !V || !W || !incr_rt || (c->g.np1 == 2 && !incr_r))
    return fail(c, ASAM_E_INVALID, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  double *db, *dU, *dV, *dW, *dit, *dir;
  int rc = pj_stage(c, &db, &dU, &dV, &dW, &dit, &dir);
  if (rc) return rc;
  const Geom& g = c->g;
  const long nfu = (long)(g.nx + 1) * g.ny * g.nz, nfv = (long)g.nx * (g.ny + 1) * g.nz,
             nfw = (long)g.nx * g.ny * (g.nz + 1);
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(dU, U, nfu * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dV, V, nfv * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dW, W, nfw * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dit, incr_rt, g.N * 8, hipMemcpyHostToDevice, s));
  if (incr_r) HIPCHK(c, hipMemcpyAsync(dir, incr_r, g.N * 8, hipMemcpyHostToDevice, s));
  rc = asam_sound_rhs_device(c, db, dU, dV, dW, dit, incr_r ? dir : nullptr);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(b, db, c->n * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  return 0;
}

extern "C" int asam_sound_update(asam_sound_ctx* c, double* x, double* U, double* V, double* W,
                                 double* incr_rt, double* incr_r) {
  if (!c) return ASAM_E_INVALID;
  if (!c->ready) return fail(c, ASAM_E_STATE, "update before setup");
  if (!x || !U || !V || !W || !incr_rt) return fail(c, ASAM_E_INVALID, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  double *dx, *dU, *dV, *dW, *dit, *dir;
  int rc = pj_stage(c, &dx, &dU, &dV, &dW, &dit, &dir);
  if (rc) return rc;
  const Geom& g = c->g;
  const long nfu = (long)(g.nx + 1) * g.ny * g.nz, nfv = (long)g.nx * (g.ny + 1) * g.nz,
             nfw = (long)g.nx * g.ny * (g.nz + 1);
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(dx, x, c->n * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dU, U, nfu * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dV, V, nfv * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dW, W, nfw * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dit, incr_rt, g.N * 8, hipMemcpyHostToDevice, s));
  if (incr_r) HIPCHK(c, hipMemcpyAsync(dir, incr_r, g.N * 8, hipMemcpyHostToDevice, s));
  rc = asam_sound_update_device(c, dx, dU, dV, dW, dit, incr_r ? dir : nullptr);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(x, dx, c->n * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(U, dU, nfu * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(V, dV, nfv * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(W, dW, nfw * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(incr_rt, dit, g.N * 8, hipMemcpyDeviceToHost, s));
  if (incr_r) HIPCHK(c, hipMemcpyAsync(incr_r, dir, g.N * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  return 0;
}

extern "C" int asam_dev_alloc(asam_sound_ctx* c, void** ptr, size_t bytes) {
  if (!c || !ptr) return ASAM_E_INVALID;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMalloc(ptr, bytes ? bytes : 1));
  return 0;
}
extern "C" int asam_dev_free(asam_sound_ctx* c, void* ptr) {
  if (!c) return ASAM_E_INVALID;
  if (ptr) HIPCHK(c, hipFree(ptr));
  return 0;
}
extern "C" int asam_dev_upload(asam_sound_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || !dst || !src) return ASAM_E_INVALID;
  HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
extern "C" int asam_dev_download(asam_sound_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || !dst || !src) return ASAM_E_INVALID;
  HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
extern "C" int asam_dev_memset(asam_sound_ctx* c, void* dst, int value, size_t bytes) {
  if (!c || !dst) return ASAM_E_INVALID;
  HIPCHK(c, hipMemsetAsync(dst, value, bytes, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
extern "C" int asam_dev_axpby(asam_sound_ctx* c, size_t n, double a, const double* x, double b,
                              double* y) {
  if (!c || !x || !y) return ASAM_E_INVALID;
  launch_axpby(c->stream, (long)n, a, x, b, y);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
extern "C" int asam_dev_dot(asam_sound_ctx* c, size_t n, const double* x, const double* y,
                            double* dot) {
  if (!c || !x || !y || !dot) return ASAM_E_INVALID;
  launch_dot(c->stream, (long)n, x, y, c->partial);
  launch_finalize(c->stream, FIN_SUMS, 1, blas_blocks((long)n), c->partial, c->st, 0, 0, nullptr);
  int rc = sync_state(c);
  if (rc) return rc;
  *dot = c->st_host->sum[0];
  return 0;
}
extern "C" int asam_dev_mass(asam_sound_ctx* c, double* y, const double* x) {
  if (!c || !x || !y) return ASAM_E_INVALID;
  if (!c->ready) return fail(c, ASAM_E_STATE, "before setup");
  launch_mass(c->stream, c->g, c->c64.vol, y, x);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// ------------------------------------------------------------------ measurement
extern "C" int asam_sound_time_kernel(asam_sound_ctx* c, int which, int fp32, int reps,
                                      double* ms) {
  if (!c || !ms || reps < 1 || which < 0 || which > 4) return ASAM_E_INVALID;
  if (!c->ready) return fail(c, ASAM_E_STATE, "before setup");
  HIPCHK(c, hipSetDevice(c->device));
  Enq e(c, 0);
  const int prec = fp32 ? 1 : 0;
  void* x = fp32 ? (void*)c->p32 : (void*)c->p;
  void* y = fp32 ? (void*)c->v32 : (void*)c->v;
  launch_fill(c->stream, c->n, prec, x, 1.0);
  launch_fill(c->stream, c->n, prec, y, 0.5);
  launch_fill(c->stream, c->n, prec, fp32 ? (void*)c->r32 : (void*)c->r, 0.25);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double *px = nullptr, *pU = nullptr, *pV = nullptr, *pW = nullptr, *pit = nullptr,
         *pir = nullptr;
  if (which == 3 || which == 4) {  // FP64 only; staging block zero-filled (values stay 0)
    int rc = pj_stage(c, &px, &pU, &pV, &pW, &pit, &pir);
    if (rc) return rc;
    launch_fill(c->stream, (long)(pir + c->g.N - px), 0, px, 0.0);
  }
  auto once = [&]() {
    if (which == 0) e.apply(prec, prec, prec, prec, x, y, nullptr, nullptr, nullptr, 0, nullptr);
    else if (which == 1) e.line(prec, LINE_PLAIN, x, nullptr, nullptr, y, nullptr, nullptr);
    else if (which == 3) launch_rhs(c->stream, c->g, proj_inputs(c), px, pU, pV, pW, pit, pir);
    else if (which == 4) launch_update(c->stream, c->g, proj_inputs(c), px, pU, pV, pW, pit, pir);
    else {
      memset(c->st_host, 0, sizeof(KState));
      launch_bicg_update(c->stream, c->n, prec, x, y, y, fp32 ? (void*)c->r32 : (void*)c->r, y,
                         y, c->st, c->partial);
    }
  };
  if (which == 2) {
    memset(c->st_host, 0, sizeof(KState));
    int rc = upload_state(c);
    if (rc) return rc;
  }
  once();
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  for (int q = 0; q < reps; ++q) once();
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float t = 0.f;
  HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
  HIPCHK(c, hipGetLastError());
  *ms = (double)t / reps;
  return 0;
}

// Launch-shape autotuner.  Objective: device time of a fixed number of BiCGStab
// iterations (every kernel and finalize of the iteration, in the order a solve runs
// them, on this context's operator) in the precision the control selects.  Timing whole
// iterations rather than isolated launches keeps each kernel's cache context the same as
// in a solve.  Coordinate search: apply shapes first, then line shapes.
template <typename VT>
static int tune_iters(asam_sound_ctx* c, int prec, int iters, double* ms) {
  VT *x, *r, *rt, *p, *ph, *v, *rh, *t;
  if (prec) {
    x = (VT*)c->d32; r = (VT*)c->r32; rt = (VT*)c->rt32; p = (VT*)c->p32; ph = (VT*)c->ph32;
    v = (VT*)c->v32; rh = (VT*)c->rh32; t = (VT*)c->t32;
  } else {
    x = (VT*)c->xd; r = (VT*)c->r; rt = (VT*)c->rt; p = (VT*)c->p; ph = (VT*)c->ph;
    v = (VT*)c->v; rh = (VT*)c->rh; t = (VT*)c->t;
  }
  launch_fill(c->stream, c->n, prec, x, 0.0);
  launch_fill(c->stream, c->n, prec, r, 1.0);
  launch_fill(c->stream, c->n, prec, rt, 1.0);
  KState* hs = c->st_host;
  memset(hs, 0, sizeof(KState));
  hs->iter = 1; hs->maxiter = 1 << 30; hs->tol = 0.0; hs->rho = 1.0; hs->rnorm0 = 1.0;
  int rc = upload_state(c);
  if (rc) return rc;
  Enq e(c, 0);
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  for (int k = 0; k < iters; ++k) bicg_iteration<VT>(e, prec, x, r, rt, p, ph, v, rh, t);
  e.flush();
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  HIPCHK(c, hipGetLastError());
  float tm = 0.f;
  HIPCHK(c, hipEventElapsedTime(&tm, c->ev0, c->ev1));
  *ms = (double)tm / iters;
  return 0;
}

extern "C" int asam_sound_autotune(asam_sound_ctx* c, int reps, asam_sound_tuning* best,
                                   double* ms) {
  if (!c || reps < 1 || reps > 1000) return ASAM_E_INVALID;
  if (!c->ready) return fail(c, ASAM_E_STATE, "autotune before setup");
  HIPCHK(c, hipSetDevice(c->device));
  const int prec = c->ctl.precision == ASAM_PREC_MIXED ? 1 : 0;
  auto measure = [&](double* out) -> int {
    double m = 0.0;
    int rc = prec ? tune_iters<float>(c, 1, 2, &m) : tune_iters<double>(c, 0, 2, &m);  // warm
    if (rc) return rc;
    rc = prec ? tune_iters<float>(c, 1, 2 * reps, &m) : tune_iters<double>(c, 0, 2 * reps, &m);
    *out = m;
    return rc;
  };
  const Tuning t0 = c->g.tune;
  Tuning bt = t0;
  double best_ms = 1e30;
  const int variants[] = {0, 1}, bxs[] = {64, 128, 256}, xcds[] = {1, 0},
            wgs[] = {2048, 4096, 8192, 16384};
  for (int var : variants)
    for (int bx : bxs)
      for (int xcd : xcds)
        for (int wg : wgs) {
          c->g.tune = bt;
          c->g.tune.apply_variant = var; c->g.tune.apply_bx = bx;
          c->g.tune.apply_xcd = xcd; c->g.tune.apply_wg = wg;
          if (apply_blocks(c->g, 0) > kRedBlocks || apply_blocks(c->g, 1) > kRedBlocks) continue;
          double m = 0;
          int rc = measure(&m);
          if (rc) { c->g.tune = t0; return rc; }
          if (m < best_ms) { best_ms = m; bt = c->g.tune; }
        }
  // band-ordered variant: width, rows per band (0: one band per XCD), workgroups
  const int rows[] = {0, 16, 32, 64}, bwgs[] = {1024, 2048, 4096};
  for (int bx : bxs)
    for (int rb : rows)
      for (int wg : bwgs) {
        c->g.tune = bt;
        c->g.tune.apply_variant = 2; c->g.tune.apply_bx = bx;
        c->g.tune.apply_band_rows = rb; c->g.tune.apply_wg = wg;
        double m = 0;
        int rc = measure(&m);
        if (rc) { c->g.tune = t0; return rc; }
        if (m < best_ms) { best_ms = m; bt = c->g.tune; }
      }
  // row-blocked z-march: rows per thread, width, XCD order, workgroups
  // width 32: a wavefront spans two stacked row blocks, so the rows between them are
  // loaded by one wavefront in one iteration
  // (fewer workgroups than the other variants: longer z-marches per tile; the best count
  // depends on the grid and is not monotonic)
  // (loads two levels ahead: FP32 2x2 blocks only)
  const int rbs[] = {2, 4}, rbxs[] = {32, 64, 128, 256}, rwgs[] = {256, 512, 1024, 2048, 4096};
  const int depths = (prec == 1 && c->g.np1 == 2) ? 2 : 1;
  for (int dp = 1; dp <= depths; ++dp)
  for (int rr : rbs)
    for (int bx : rbxs)
      for (int xcd : xcds)
        for (int wg : rwgs) {
          c->g.tune = bt;
          c->g.tune.apply_variant = 3; c->g.tune.apply_rows = rr; c->g.tune.apply_bx = bx;
          c->g.tune.apply_xcd = xcd; c->g.tune.apply_wg = wg; c->g.tune.apply_depth = dp;
          if (dp == 2 && rr == 4) continue;  // two rows only
          if (apply_blocks(c->g, 0) > kRedBlocks || apply_blocks(c->g, 1) > kRedBlocks) continue;
          double m = 0;
          int rc = measure(&m);
          if (rc) { c->g.tune = t0; return rc; }
          if (m < best_ms) { best_ms = m; bt = c->g.tune; }
        }
  // budgets past 96 KiB stage every level of an FP64 2x2 column of up to 96 levels; 76 KiB
  // is the most that still lets two workgroups share a CU's 160 KiB
  const int windows[] = {1, 4, 8}, ldss[] = {1, 0, 16, 32, 48, 76, 80, 100, 128};
  const Tuning ba = bt;
  for (int w : windows)
    for (int l : ldss) {
      c->g.tune = ba;
      c->g.tune.line_window = w; c->g.tune.line_lds = l;
      double m = 0;
      int rc = measure(&m);
      if (rc) { c->g.tune = t0; return rc; }
      if (m < best_ms) { best_ms = m; bt = c->g.tune; }
    }
  // cache policy of the line solve's factor loads
  const Tuning bn = bt;
  for (int ntl : {0, 1}) {
    c->g.tune = bn;
    c->g.tune.line_nt = ntl;
    double m = 0;
    int rc = measure(&m);
    if (rc) { c->g.tune = t0; return rc; }
    if (m < best_ms) { best_ms = m; bt = c->g.tune; }
  }
  // scalar tails as finalize launches, in the producing kernels' last workgroups, or
  // folded into the consuming kernels
  const Tuning bl = bt;
  for (int ft : {0, 1, 2}) {
    c->g.tune = bl;
    c->g.tune.fused_tail = ft;
    double m = 0;
    int rc = measure(&m);
    if (rc) { c->g.tune = t0; return rc; }
    if (m < best_ms) { best_ms = m; bt = c->g.tune; }
  }
  c->g.tune = bt;
  memset(c->st_host, 0, sizeof(KState));
  if (best) asam_sound_get_tuning(c, best);
  if (ms) { ms[0] = best_ms; ms[1] = 0.0; }
  return 0;
}

extern "C" int asam_sound_profile_enable(asam_sound_ctx* c, int on) {
  if (!c) return ASAM_E_INVALID;
  c->prof_on = on ? 1 : 0;
  return 0;
}

extern "C" int asam_sound_profile_read(asam_sound_ctx* c, double ms[8], double bytes[8],
                                       int launches[8]) {
  if (!c || !ms || !bytes || !launches) return ASAM_E_INVALID;
  for (int q = 0; q < 8; ++q) {
    ms[q] = c->prof_ms[q]; bytes[q] = c->prof_bytes[q]; launches[q] = c->prof_launches[q];
    c->prof_ms[q] = 0; c->prof_bytes[q] = 0; c->prof_launches[q] = 0;
  }
  return 0;
}

extern "C" int asam_sound_time_copy(asam_sound_ctx* c, size_t bytes, int reps, double* ms) {
  if (!c || !ms || reps < 1 || bytes < 16) return ASAM_E_INVALID;
  HIPCHK(c, hipSetDevice(c->device));
  void *a = nullptr, *b = nullptr;
  HIPCHK(c, hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); return fail(c, ASAM_E_NOMEM, "copy"); }
  const long n16 = (long)(bytes / 16);
  (void)hipMemsetAsync(a, 0, bytes, c->stream);
  launch_copy16(c->stream, n16, a, b);
  (void)hipEventRecord(c->ev0, c->stream);
  for (int q = 0; q < reps; ++q) launch_copy16(c->stream, n16, a, b);
  (void)hipEventRecord(c->ev1, c->stream);
  hipError_t e = hipEventSynchronize(c->ev1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, c->ev0, c->ev1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return fail(c, ASAM_E_HIP, hipGetErrorString(e));
  *ms = (double)t / reps;
  return 0;
}

extern "C" int asam_sound_last_counts(const asam_sound_ctx* c, int counts[8]) {
  if (!c || !counts) return ASAM_E_INVALID;
  for (int q = 0; q < 8; ++q) counts[q] = c->counts[q];
  return 0;
}
