// internal.h -- device data layout and kernel launchers of the gfx950 sound solver.
//
// Data layout in HBM (one block, N = nx*ny*nz cells, x fastest):
//   stencil coefficients, one word per cell, structure-of-arrays:
//     MultiTriT  (np1=1): kx,ky,kz (face coupling incl. face theta), vol, dut     = 5 words
//     MultiTriTB (np1=2): kx,ky,kz (coupling without theta), tx,ty,tz (face theta),
//                         gz (gravity coupling of the z-face above), vol, dut     = 9 words
//     The face (c, c+e) of direction e is stored at cell c; wall faces hold 0.
//   line factors (vertical-line LU of the z-tridiagonal): pl, pt, pw,
//     np1*np1 words each per cell (2x2 blocks column-major [11,21,12,22]).
//   vectors: p(1:np1, nx, ny, nz) as in the reference, component fastest.
// Every array exists in FP64 and (after set-up) in FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace asam {

constexpr double kEps = 1.0e-30;          // Kind_Mod Eps
constexpr double kRd = 1004.0 - 717.0;    // Parameter_Mod Rd = Cpd - Cvd
constexpr double kCpd = 1004.0;
constexpr double kRv = 461.0;             // Parameter_Mod Rv, Cpv, Cpl (vapour, liquid water)
constexpr double kCpv = 1885.0;
constexpr double kCpl = 4186.0;
constexpr int kMaxM = 32;                 // largest (F)GMRES restart length
constexpr int kRedBlocks = 16384;         // partial-sum slots per reduction quantity

template <typename T>
struct Coef {
  T *kx = nullptr, *ky = nullptr, *kz = nullptr;
  T *tx = nullptr, *ty = nullptr, *tz = nullptr, *gz = nullptr;
  T *vol = nullptr, *dut = nullptr;
  T *pl = nullptr, *pt = nullptr, *pw = nullptr;
  T *pkz = nullptr;  // FP32 2x2 blocks: (kz, tz, gz, vol) of each cell
};

// Launch-shape knobs (set by the autotuner or asam_sound_set_tuning).
struct Tuning {
  int apply_variant = 3;  // 0: z-march (column kept in registers), 1: one cell per thread,
                          // 2: one cell per thread in band order (each XCD its own bands),
                          // 3: z-march of apply_rows rows per thread
  int apply_wg = 512;     // target workgroups per apply launch
  int line_window = 0;    // levels per load window of the line solve (1, 4, 8; 0: by type)
  int line_lds = 1;       // LDS staging budget of the line solve's forward sweep per
                          // workgroup: 0 none (output vector), 1 default, n >= 2 n KiB
  int apply_xcd = 1;      // remap workgroups so each XCD gets a contiguous range of tiles
  int apply_bx = 128;     // apply workgroup width along x (32, 64, 128 or 256; 256 threads)
  int line_tiled = 1;     // line factors stored column-tiled (64 columns x nz levels)
  int apply_band_rows = 0;  // band-ordered apply: rows per band (0: ny/8)
  int apply_rows = 2;     // variant 3: rows per thread (2 or 4)
  int fused_tail = 0;     // 1: BiCGStab and refinement tails run by the producing kernel's
                          // last workgroup; 2: the BiCGStab iteration's tails folded into
                          // the kernels that consume their scalars (0: a separate finalize
                          // launch each)
  int apply_depth = 1;    // variant 3, FP32 2x2 blocks, two rows: levels of loads ahead (1, 2)
  int line_nt = 2;        // line solve's factor loads (and the vectors it reads once):
                          // 0 plain, 1 nontemporal, 2 nontemporal when the factors of the
                          // solve's precision exceed kLineNtBytes
};

// Factor footprint above which the line solve's loads default to nontemporal: below it the
// factors stay resident in the 256 MB Infinity Cache across solves (measured: FP32 MultiTriT
// 256x256x64, 50 MB of factors, line 24.8 us plain against 28.6 us nontemporal; FP32
// MultiTriTB, 201 MB, 64.5 us plain against 58.7 us nontemporal).
constexpr long kLineNtBytes = 64l << 20;

struct Geom {
  int nx, ny, nz, np1;
  long N;      // cells
  double fa;   // FacAnela
  Tuning tune;
};

// Index of the line-factor entry of cell (i,j,k).  Column-tiled layout: 64 consecutive
// x-columns of one y-row form a tile whose nz levels are contiguous, so a wavefront
// (one tile) streams one contiguous block per array through its sweeps.
__host__ __device__ inline long fac_index(const Geom& g, int i, int j, int k) {
  if (g.tune.line_tiled) {
    const long gxl = (g.nx + 63) / 64;
    return (((long)j * gxl + (i >> 6)) * g.nz + k) * 64 + (i & 63);
  }
  return i + (long)g.nx * (j + (long)g.ny * k);
}
// Words per cell of the line solve's lower blocks L_k.  A 2x2 L_k couples level k to the
// cell below through one face: L_k = (tz, 1)^T (-kz DUT, gz V) of that cell, rank one, so
// it is stored as the three words (tz, -kz DUT, gz V).
constexpr int lower_words(int np1) { return np1 == 2 ? 3 : 1; }
// Entries per line-factor array (tiled layout pads x to a multiple of 64).
inline long fac_cells(const Geom& g) { return ((long)(g.nx + 63) / 64) * 64 * g.ny * g.nz; }

// Device-resident scalar state of one Krylov solve (written only by the
// single-thread tails of the finalize kernels, read by every kernel).
struct KState {
  // BiCGStab
  double rho, rho_old, alpha, omega, beta;
  double rnorm0, rnorm, tol;
  double sum[4];
  int iter, maxiter, done, status;
  int half, result, first, pad;
  // (F)GMRES scalars
  double gbeta, grho, grho0, gtol;
  int m, step, nR, cycle_done;
  int total, cycles, maxcycles, gstatus;
  // mixed refinement
  double outer_r0, outer_r, outer_tol, inner_tol;
  int outer, inner_total, outer_done, pad2;
  int xpar;             // ensemble refinement: which of the two x buffers holds x
  int pad3;
  int xnz, xz;          // initial guess has a nonzero entry / is all zero (k_any_nonzero)
  int pad4, pad5;
  int hist[16];         // inner iterations of refinement step q (q < 16)
  // (F)GMRES arrays, last: the tails that do not use them copy only the part above
  double H[(kMaxM + 1) * kMaxM];  // H[row*kMaxM + col]
  double cs[kMaxM], sn[kMaxM], bhat[kMaxM + 1], y[kMaxM], vscale[kMaxM + 1];
};

// Scalar tail of a reduction run inside a kernel in place of a finalize launch (kind < 0:
// none): by the last workgroup of the kernel that produced the partial sums (fused tail),
// or by every workgroup of the kernel that needs the scalars (folded tail).  The kinds are
// those whose step reads and writes only the state's scalars ahead of H.  counter (fused
// tails only) counts the arriving workgroups; it is zero between launches (the last
// workgroup resets it).
struct KTail {
  int kind = -1;
  int nq = 0, nblk = 0, nblk_hi = 0;  // slots, partials of slots 0-1 and of slots 2-3
  int qmask = 0xF;                    // slots summed (the others enter the step as 0)
  const double* partial = nullptr;
  KState* st = nullptr;
  unsigned int* counter = nullptr;
  // folded tails: the state is read from src (null: st) and also written to copy
  const KState* src = nullptr;
  KState* copy = nullptr;
};

// Partial-sum buffer: two regions of 4 slots of kRedBlocks.  With folded tails the
// BiCGStab update writes its sums to the second region, because the update's own
// workgroups still read the first while they fold the previous tail.
constexpr long kPartialRegion = 4L * kRedBlocks;

// ------------------------------------------------------------------ set-up
void launch_weights(hipStream_t s, const Geom& g, const double* rho, const double* rth,
                    const double* T, const double* rhov, const double* rhol,
                    const double* rhor, double* theta, double* dut);
void launch_face_theta(hipStream_t s, const Geom& g, const double* vol, const double* theta,
                       double* thu, double* thv, double* thw);
// Face coefficients from areas, volumes, spacings and face theta (full face shapes).
void launch_coef(hipStream_t s, const Geom& g, int grad_full, double beta0_dt, double grav,
                 const double* FU, const double* FV, const double* FW, const double* vol,
                 const double* dx, const double* dy, const double* dz, const double* thu,
                 const double* thv, const double* thw, Coef<double>& c);
void launch_factor(hipStream_t s, const Geom& g, Coef<double>& c);
void launch_convert(hipStream_t s, const Geom& g, const Coef<double>& c64, Coef<float>& c32);

// ------------------------------------------------------------------ apply
// y = Op(x) = -A x (or -A^T x), optionally residual form y = b - Op(x_eff) with
// x_eff = x (+ dx, written back to xw).  Partial sums: q0 = y.y, q1 = y.u1.
enum { DOT_YY = 1, DOT_YU = 2 };
struct ApplyReq {
  int tr = 0;                  // transpose
  int prec = 0;                // 0 FP64 coefficients, 1 FP32
  int vx = 0, vo = 0, vu = 0;  // vector element types: 0 double, 1 float
  const void* x = nullptr;
  const float* dx = nullptr;   // mixed refinement correction (x must be double)
  double* xw = nullptr;
  const double* b = nullptr;   // residual form
  void* y = nullptr;
  void* y2 = nullptr;
  const void* u = nullptr;
  int dots = 0;
  double* partial = nullptr;
  const int* done = nullptr;
  // x known to be zero (x unused): the residual form is y = b, zero on the cells this
  // mask marks dead (their operator row is empty), the same values as the full apply
  const unsigned char* x0_dead = nullptr;
  KTail tail;                  // fused scalar tail of the dots (needs dots)
};
// Workgroups (= partial sums) of an apply launch; fp32: FP32 coefficients.
int apply_blocks(const Geom& g, int fp32);
// dead[c] = 1 where the residual form writes 0 (no volume and no face coefficient), from
// the FP64 coefficients with the apply's own test.
void launch_dead_mask(hipStream_t s, const Geom& g, const Coef<double>& c64, unsigned char* dead);
// Returns 0, or ASAM_E_INVALID (nothing launched) for an option set no kernel is built for.
int launch_apply(hipStream_t s, const Geom& g, const Coef<double>& c64, const Coef<float>& c32,
                 const ApplyReq& r);

// ------------------------------------------------------------------ line solve
// out = -M^-1 in (PSolve1 semantics) with in formed on the fly:
//   LINE_PLAIN:  in = src * (*scale or 1)
//   LINE_BICG_P: in = first ? r : r + beta*(p - omega*v); p := in
//   LINE_BICG_S: in = r - alpha*v; r := in; partial q0 = in.in
enum { LINE_PLAIN = 0, LINE_BICG_P = 1, LINE_BICG_S = 2 };
struct LineReq {
  int tr = 0, prec = 0, vt = 0;  // vt: vector type 0 double 1 float
  int mode = LINE_PLAIN;
  const void* src = nullptr;     // PLAIN source / r
  void* p = nullptr;             // BICG_P: p (in/out); BICG_S: r (in/out)
  const void* v = nullptr;
  void* out = nullptr;
  const double* scale = nullptr;
  const KState* st = nullptr;
  double* partial = nullptr;
  const int* done = nullptr;
  int use_lds = 1;
  int qslot = 0;  // partial-sum quantity slot of the LINE_BICG_S norm
  KTail fold;     // tail folded into the launch: every workgroup runs the scalar step first
};
int line_blocks(const Geom& g);
void launch_line(hipStream_t s, const Geom& g, const Coef<double>& c64, const Coef<float>& c32,
                 const LineReq& r);

// ------------------------------------------------------------------ BLAS-1 + reductions
int blas_blocks(long n);
// *xnz = 1 and *xz = 0 if x (n doubles) has an entry != 0; both untouched otherwise.
void launch_any_nonzero(hipStream_t s, long n, const double* x, int* xnz, int* xz);
// BiCGStab tail: x += alpha pHat + omega rHat; r -= omega t; q0 = r.r, q1 = rt.r.
// With st->half set: x += alpha pHat only.
// fold: the half-step/omega tail of the preceding stencil apply, folded into the launch.
void launch_bicg_update(hipStream_t s, long n, int vt, void* x, const void* ph, const void* rh,
                        void* r, const void* t, const void* rt, const KState* st,
                        double* partial, const KTail& tail = KTail(),
                        const KTail& fold = KTail());
// (F)GMRES modified Gram-Schmidt step: w -= H[k-1][i]*vscale[k-1]*v_{k-1} (k>0); then
// q0 = w.v_k (k<=i) or w.w (k==i+1), into partial slot 0 of `partial`.  hsrc: null (H[k-1][i]
// is in the state), or the previous step's partial products, whose sum is H[k-1][i]
// (recorded in the state by workgroup 0).
void launch_mgs(hipStream_t s, long n, int vt, void* w, const void* vprev, const void* vk,
                int k, int i, KState* st, double* partial, int chk_done = 0,
                const double* hsrc = nullptr);
// x += sum_{q<nR} y_q * Z_q  (Z_q = z vectors, FGMRES) with optional scale vscale (GMRES).
void launch_gmres_update(hipStream_t s, long n, int vt, void* x, void* const* zs, int m,
                         int use_vscale, const KState* st);
// FGMRESV classical Gram-Schmidt of one Arnoldi step (Iter_Mod.f90:238-248): dots of the
// new vector w with up to 4 basis vectors k0.. (partial slots 0..nq-1), and the removal of
// all i+1 projections from w with |w|^2 (slot 0).  Basis vector k is vs[k] * vscale[k].
void launch_cgs_dots(hipStream_t s, long n, int vt, const void* w, void* const* vs, int k0,
                     int nq, const KState* st, double* partial);
void launch_cgs_update(hipStream_t s, long n, int vt, void* w, void* const* vs, int i,
                       const KState* st, double* partial);
void launch_dot(hipStream_t s, long n, const double* x, const double* y, double* partial);
void launch_axpby(hipStream_t s, long n, double a, const double* x, double b, double* y);
void launch_fill(hipStream_t s, long n, int vt, void* x, double v);
void launch_copy_cvt(hipStream_t s, long n, int vt_dst, void* dst, int vt_src, const void* src);
void launch_mass(hipStream_t s, const Geom& g, const double* vol, double* y, const double* x);
void launch_theta_vjp(hipStream_t s, const Geom& g, const Coef<double>& c,
                      const double* thfaces, double* grad, const double* mu, const double* x);

// Finalize kernels: sum partial slots (deterministic order) then run the scalar logic.
enum FinKind {
  FIN_SUMS = 0,         // sums only -> st->sum[]
  FIN_BICG_INIT,        // rnorm0, rho, iter=1
  FIN_BICG_ALPHA,       // alpha = rho / (rt.v)
  FIN_BICG_HALF,        // rnorm = |s|; half-step exit test
  FIN_BICG_OMEGA,       // omega = (t.s)/(t.t)
  FIN_BICG_HALF_OMEGA,  // half-step exit test on |s| (slot 2), else omega (slots 0, 1)
  FIN_BICG_END,         // rnorm = |r|; exit tests; rho, beta for the next iteration
  FIN_GM_INIT,          // beta = |v0|; restart / exit bookkeeping
  FIN_GM_H,             // H[k][i] = sum   (k <= i)
  FIN_GM_HV,            // H[k+q][i] = sum q, q < nq (classical Gram-Schmidt pass)
  FIN_GM_LAST,          // H[i+1][i] = |w|, Givens, convergence test
  FIN_DOT,              // store sum[0] to host-visible double (out)
  FIN_MIXED_OUTER,      // FP64 refinement residual -> next FP32 inner solve or exit
  FIN_MIXED_GM,         // FP64 refinement residual -> one FP32 (F)GMRES cycle or exit
};
// Quantities q < 2 were written by nblk workgroups, q >= 2 by nblk_hi workgroups.
void launch_finalize(hipStream_t s, int kind, int nq, int nblk, const double* partial,
                     KState* st, int k, int i, double* out, int nblk_hi = 0);
// Back-substitution of the (F)GMRES least-squares problem (single thread).
void launch_gm_backsolve(hipStream_t s, KState* st);

void launch_copy16(hipStream_t s, long n16, const void* a, void* b);

// ------------------------------------------------------------------ either side of the solve
// Geometry and weights the right-hand side and the post-solve update need (device).
struct ProjInputs {
  int grad_full;
  double beta0_dt, grav;
  const double *FU, *FV, *FW;     // face areas, shapes (nx+1)ny nz, nx(ny+1)nz, nx ny(nz+1)
  const double *thu, *thv, *thw;  // face theta, same shapes
  const double *V, *dut, *dx, *dy, *dz;
};
// b = bd * div(F theta_f u) - V incr_rt (and TB: bd * div(F u) - V incr_r)
void launch_rhs(hipStream_t s, const Geom& g, const ProjInputs& in, double* b, const double* U,
                const double* V, const double* W, const double* incr_rt, const double* incr_r);
// face velocities -= grad(DUT x1) (+ TB buoyancy), increments -= bd div(...)/V, x1 *= DUT,
// x2 *= DUR (= 0); incr_r may be null
void launch_update(hipStream_t s, const Geom& g, const ProjInputs& in, double* x, double* U,
                   double* V, double* W, double* incr_rt, double* incr_r);

}  // namespace asam
