/*
 * asam_sound.h -- C-ABI of the MI355X (gfx950) sound stage solver.
 *
 * Drop-in for ASAM's implicit sound/buoyancy stage solve: the system built by
 * JacAccGrav (reference SRC/MULTI/JacAccGrav_Mod.f90:304-471) and solved by
 * SolveSound (reference SRC/SOLU/Iter_Mod.f90:9-30), with the reference's
 * conventions kept at this boundary:
 *   - the Krylov operator is MatVec1 = -A x (reference SRC/MATRIX/MatVec_Mod.f90:232),
 *     so SolveSound solves (-A) x = b;
 *   - the preconditioner has PSolve1 semantics, -M^-1 x (reference
 *     SRC/MATRIX/PreCond_Mod.f90:1555-1571), M the vertical-line block of A;
 *   - vectors are the `p` component of TYPE(PressureVelocity)
 *     (reference SRC/MODULE/VecPrVel_Mod.f90:11-18): p(1:np1, nx, ny, nz),
 *     component fastest, then x, y, z (Fortran column-major);
 *   - geometry and state arrays use the reference's interior/face shapes:
 *     FU(0:nx,1:ny,1:nz), FV(1:nx,0:ny,1:nz), FW(1:nx,1:ny,0:nz), cells (nx,ny,nz).
 *
 * All functions return 0 on success, a negative ASAM_E_* code on error; they
 * never abort the host.  A context is bound to the HIP device that was
 * current when it was created and owns one HIP stream; calls on one context
 * are not re-entrant, calls on different contexts are independent.
 */
#ifndef ASAM_SOUND_H
#define ASAM_SOUND_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASAM_OK 0
#define ASAM_E_INVALID (-1)   /* bad argument / size */
#define ASAM_E_HIP (-2)       /* HIP runtime error */
#define ASAM_E_STATE (-3)     /* call order (e.g. solve before setup) */
#define ASAM_E_NOMEM (-4)     /* device allocation failed */
#define ASAM_E_NAN (-5)       /* non-finite input or iterate */

/* Krylov method (MethSound, reference SRC/MODULE/Control_Mod.f90:64). */
#define ASAM_METHOD_BICGSTAB 0 /* SpMatrix4CellPar_Mod.f90:245-348 sequence */
#define ASAM_METHOD_FGMRES 1   /* Iter_Mod.f90:121-199 (reference default) */
#define ASAM_METHOD_GMRES 2    /* Iter_Mod.f90:32-119 (left preconditioned) */
#define ASAM_METHOD_FGMRESV 3  /* Iter_Mod.f90:201-288 (FGMRES, classical Gram-Schmidt) */

/* Arithmetic of the solve. */
#define ASAM_PREC_FP64 0  /* plain FP64 drop-in */
#define ASAM_PREC_MIXED 1 /* FP32 operator/preconditioner/inner Krylov; FP64 residual,
                             dots and refinement (BiCGStab, FGMRES or FGMRESV) */

/* Solve status (asam_sound_info.status). */
#define ASAM_CONVERGED 0
#define ASAM_NOT_CONVERGED 1
#define ASAM_BREAKDOWN 2 /* rho == 0 in BiCGStab (reference returns Tol = 0) */

typedef struct asam_sound_ctx asam_sound_ctx;

/* Grid and physical constants of one block (one stage system). */
typedef struct {
  int nx, ny, nz;    /* interior cells */
  int np1;           /* 1: MultiTriT (rho*theta'), 2: MultiTriTB (rho*theta', rho') */
  int grad_full;     /* Control_Mod GradFull (reference default 1) */
  double beta0_dt;   /* beta0*dtP: Rosenbrock gamma times the step */
  double grav;       /* GravComp: Grav for MultiTriTB, 0 otherwise */
  double fac_anela;  /* Physics_Mod FacAnela: 1, 0 when anelastic */
} asam_sound_grid;

/* Solver controls (reference &ModelControl knobs). */
typedef struct {
  int method;        /* ASAM_METHOD_* */
  int step_gmres;    /* restart length m (StepGMRES, default 10), <= 32 */
  int precision;     /* ASAM_PREC_* */
  double inner_tol;  /* mixed: relative reduction asked of each FP32 inner solve */
  int max_inner;     /* mixed: cap on inner iterations per refinement step */
  int max_outer;     /* mixed: cap on refinement steps */
} asam_sound_control;

/* Outcome of one solve. */
typedef struct {
  int iterations;         /* BiCGStab: completed iterations (a half-step exit counts as one);
                             (F)GMRES: Arnoldi steps summed over restarts;
                             mixed: inner iterations summed over refinement steps */
  int restarts;           /* (F)GMRES restart cycles / mixed: refinement steps run (FP32
                             inner solves or cycles) */
  int status;             /* ASAM_CONVERGED / ASAM_NOT_CONVERGED / ASAM_BREAKDOWN */
  int reserved;
  double rel_residual;    /* final ||r|| / ||r0|| as tracked by the method */
  double abs_residual;    /* final ||r|| as tracked by the method */
  double ms;              /* device time of the solve (HIP events), milliseconds */
} asam_sound_info;

/* Launch-shape knobs of the kernels (no effect on results beyond round-off order of
 * the reductions; the defaults are tuned for MI355X). */
typedef struct {
  int apply_variant; /* 0: z-march keeping the column in registers, 1: one cell per thread,
                        2: one cell per thread in band order (each XCD streams its own bands
                        of rows plane by plane), 3: z-march of apply_rows rows of a column
                        per thread (the inner rows' y-neighbours come from registers) */
  int apply_wg;      /* target workgroups per stencil-apply launch (256..16384; variant 2:
                        the workgroups of the launch, each looping over its XCD's tiles) */
  int line_window;   /* levels per load window of the line solve: 1, 4 or 8 */
  int line_lds;      /* LDS budget per workgroup for the line solve's forward sweep: 0 none
                        (staged in the output vector), 1 default (60 KiB), 2..160 that
                        many KiB; the lowest levels that fit are staged in LDS, the rest
                        in the output vector */
  int fused_tail;    /* 1: the scalar steps of BiCGStab and of the mixed refinement run
                        in the last workgroup of the kernel that produced their sums;
                        2: the three scalar steps of a BiCGStab iteration run in every
                        workgroup of the kernel that needs their scalars (alpha in the
                        s-step line solve, omega in the update, the exit test and beta in
                        the next p-step line solve); 0 (default): a separate finalize
                        launch each.  Same bits in all three; the autotuner times them.
                        (F)GMRES tails are always separate launches */
  int apply_xcd;     /* 1: contiguous tile ranges per XCD, 0: round-robin order */
  int apply_bx;      /* apply workgroup width along x: 32, 64, 128 or 256 (256 threads; at 32
                        a wavefront covers two stacked row blocks of variant 3) */
  int line_tiled;    /* 1: line factors stored in 64-column tiles, levels contiguous */
  int apply_band_rows; /* variant 2: rows per band (0: ny/8, one band per XCD) */
  int apply_rows;    /* variant 3: rows of a column each thread walks together, 2 or 4 */
  int line_nt;       /* cache policy of the line solve's factor loads (and of the vectors it
                        reads once): 0 plain, 1 nontemporal, 2 (default) nontemporal when the
                        factors of the solve's precision exceed 64 MiB, plain below, where
                        they stay resident in the Infinity Cache.  Same bits either way */
  int apply_depth;   /* variant 3 with FP32 coefficients, 2x2 blocks and two rows: the
                        levels a thread loads ahead of the one it computes, 1 (default;
                        the z-face coefficients and volumes then come from a packed copy,
                        16 bytes per cell) or 2.  Same bits */
} asam_sound_tuning;

/* ---- context ---------------------------------------------------------- */
int asam_sound_create(asam_sound_ctx **ctx, const asam_sound_grid *grid);
int asam_sound_destroy(asam_sound_ctx *ctx);
int asam_sound_default_control(asam_sound_control *ctl);
int asam_sound_set_control(asam_sound_ctx *ctx, const asam_sound_control *ctl);
int asam_sound_get_tuning(const asam_sound_ctx *ctx, asam_sound_tuning *t);
int asam_sound_set_tuning(asam_sound_ctx *ctx, const asam_sound_tuning *t);
/* Launch-shape autotuner: for every candidate shape of the stencil apply (variant, width,
 * XCD order, workgroup count), then of the line solve (window, LDS staging), then with
 * fused and with separate scalar tails (fused_tail), times whole
 * BiCGStab iterations (all kernels and finalizes in solve order) on this context's operator
 * in the precision the control selects (FP32 kernels for ASAM_PREC_MIXED), 2*reps
 * iterations each, and keeps the fastest.  best and ms[2] (ms per iteration of the chosen
 * shapes; ms[1] = 0) may be NULL.  Call after set-up; it uses the solve's work vectors and
 * changes later results only in the summation order of reductions. */
int asam_sound_autotune(asam_sound_ctx *ctx, int reps, asam_sound_tuning *best, double *ms);
/* Last error message of this context (static storage of the context). */
const char *asam_sound_last_error(const asam_sound_ctx *ctx);

/* ---- operator set-up (once per time step; reference JacAccGrav) -------- */
/* From geometry and dry thermodynamic state, host arrays:
 *   FU (nx+1)*ny*nz, FV nx*(ny+1)*nz, FW nx*ny*(nz+1), VolC/rho/rhotheta/T nx*ny*nz,
 *   dx nx, dy ny, dz nz.
 * Replaces JacAccGrav -> WeightDiagCompute('Density') -> LaplTCompute/LaplTBCompute
 * (JacAccGrav_Mod.f90:304-471, 4254-4274, 893-1484). */
int asam_sound_setup(asam_sound_ctx *ctx, const double *FU, const double *FV, const double *FW,
                     const double *VolC, const double *dx, const double *dy, const double *dz,
                     const double *rho, const double *rhotheta, const double *T);
/* Moist form of asam_sound_setup: vapour rhov, cloud water rhol and rain rhor densities
 * (nx*ny*nz each; any may be NULL, meaning none of that species) enter DUT as in
 * WeightDiagCompute 'Density' (JacAccGrav_Mod.f90:4254-4270): RhoD = Rho - RhoV - RhoL - RhoR,
 * Rm = Rd + Rv RhoV/RhoD, Cpml = Cpd + Cpv RhoV/RhoD + Cpl (RhoL + RhoR)/RhoD,
 * DUT = T/(rho*theta) (RhoD Rd + RhoV Rv)/(1 - Rm/Cpml).  rho is the total density. */
int asam_sound_setup_moist(asam_sound_ctx *ctx, const double *FU, const double *FV,
                           const double *FW, const double *VolC, const double *dx,
                           const double *dy, const double *dz, const double *rho,
                           const double *rhotheta, const double *T, const double *rhov,
                           const double *rhol, const double *rhor);
/* From geometry and precomputed weights DTU/DTV/DTW (face theta, face shapes) and DUT
 * (cells) -- the inputs of LaplTCompute (JacAccGrav_Mod.f90:893-1055). */
int asam_sound_setup_weights(asam_sound_ctx *ctx, const double *FU, const double *FV,
                             const double *FW, const double *VolC, const double *dx,
                             const double *dy, const double *dz, const double *DTU,
                             const double *DTV, const double *DTW, const double *DUT);

/* ---- the drop-in solve (reference SolveSound(x,b,MaxIter,Tol)) --------- */
/* Host buffers of np1*nx*ny*nz doubles.  x: initial guess in, solution out.
 * maxiter: BiCGStab iteration cap / (F)GMRES restart cycles (QMRMaxIter), in; on
 * return for BiCGStab the reference's convention (-Iter half-step exit, Iter full-step
 * exit, 2*Iter not converged, Iter with tol=0 on breakdown), unchanged otherwise.
 * tol: relative tolerance in; for BiCGStab the final ||r|| out (reference semantics).
 * With ASAM_PREC_MIXED, maxiter and tol come back unchanged and maxiter keeps its meaning
 * as a cap: BiCGStab caps every FP32 inner solve at min(maxiter, control max_inner)
 * iterations; FGMRES/FGMRESV run one restart cycle per refinement step, so maxiter caps
 * the refinement steps at min(maxiter, control max_outer).
 * A non-finite b, x0 or iterate ends the solve at once and returns ASAM_E_NAN (x then
 * holds the last iterate); the context stays usable.
 * info may be NULL. */
int asam_sound_solve(asam_sound_ctx *ctx, double *x, const double *b, int *maxiter,
                     double *tol, asam_sound_info *info);
/* Same on device buffers (allocated by asam_dev_alloc or any HIP allocation on the
 * context's device); ordered on the context's stream, synchronous on return. */
int asam_sound_solve_device(asam_sound_ctx *ctx, double *x, const double *b, int *maxiter,
                            double *tol, asam_sound_info *info);

/* ---- operator pieces (Krylov plug points) ------------------------------ */
/* y = MatVec1(x) = -A x (MatVec_Mod.f90:8-18,186-232); host buffers. */
int asam_sound_apply(asam_sound_ctx *ctx, double *y, const double *x);
/* y = -A^T x (adjoint operator; no reference counterpart). */
int asam_sound_apply_T(asam_sound_ctx *ctx, double *y, const double *x);
/* y = PSolve1(x) = -M^-1 x, M = vertical-line block of A (PreCond_Mod.f90:1555-1571). */
int asam_sound_psolve(asam_sound_ctx *ctx, double *y, const double *x);
/* y = -M^-T x. */
int asam_sound_psolve_T(asam_sound_ctx *ctx, double *y, const double *x);
/* Device-buffer forms of the four. */
int asam_sound_apply_device(asam_sound_ctx *ctx, double *y, const double *x, int transpose);
int asam_sound_psolve_device(asam_sound_ctx *ctx, double *y, const double *x, int transpose);

/* ---- adjoint solve (backward half of the forward/backward pair) -------- */
/* Solves (-A)^T lambda = g with the transposed line preconditioner, same method and
 * semantics as asam_sound_solve. */
int asam_sound_solve_T(asam_sound_ctx *ctx, double *lambda, const double *g, int *maxiter,
                       double *tol, asam_sound_info *info);
int asam_sound_solve_T_device(asam_sound_ctx *ctx, double *lambda, const double *g,
                              int *maxiter, double *tol, asam_sound_info *info);

/* Sensitivity of s = mu^T (A x) to the cell thetas (theta = rhotheta/rho) that enter A
 * through the volume-weighted face theta (CellToFaceRhoWeight, JacAccGrav_Mod.f90:3316):
 * grad[c] += d(mu^T A x)/d theta_c, device buffers (grad nx*ny*nz). */
int asam_sound_theta_vjp_device(asam_sound_ctx *ctx, double *grad, const double *mu,
                                const double *x);

/* ---- either side of the solve (one closed block, dry 'Density' weights) -- */
/* Right-hand side of the stage system (reference Rhs, MATRIX/Rhs_Mod.f90:110-157):
 *   b1 = beta0_dt * DivergenceS(F, theta_f, u) - V incr_rt     (FacAnela >= 1; else no V term)
 *   b2 = beta0_dt * Divergence(F, u) - V incr_r                 (np1 = 2)
 * U (nx+1,ny,nz), V (nx,ny+1,nz), W (nx,ny,nz+1): predicted face velocities (reference
 * Vel%uF/vF/wF); incr_rt, incr_r (nx,ny,nz): stage increments of rho*theta and rho
 * (incr_r may be NULL when np1 = 1); b: np1*nx*ny*nz.  Uses the face areas, face theta and
 * volumes of the last set-up. */
int asam_sound_rhs(asam_sound_ctx *ctx, double *b, const double *U, const double *V,
                   const double *W, const double *incr_rt, const double *incr_r);
int asam_sound_rhs_device(asam_sound_ctx *ctx, double *b, const double *U, const double *V,
                          const double *W, const double *incr_rt, const double *incr_r);
/* Post-solve update (reference UpdateVelocity, MATRIX/Rhs_Mod.f90:589-952): with x the
 * solution, x1 *= DUT (pressure); np1 = 2: interior w faces -= beta0_dt Grav
 * (V_k x2_k + V_k+1 x2_k+1)/(V_k + V_k+1 + Eps), then x2 *= DUR (= 0 for dry weights);
 * interior faces -= the pressure-gradient correction (GradFull or Dual form); then
 * incr_rt -= beta0_dt DivergenceS(u)/V and incr_r -= beta0_dt Divergence(u)/V with the
 * corrected velocities (incr_r may be NULL).  All arrays in/out; wall faces unchanged. */
int asam_sound_update(asam_sound_ctx *ctx, double *x, double *U, double *V, double *W,
                      double *incr_rt, double *incr_r);
int asam_sound_update_device(asam_sound_ctx *ctx, double *x, double *U, double *V, double *W,
                             double *incr_rt, double *incr_r);

/* ---- device memory helpers (on the context's device) ------------------- */
int asam_dev_alloc(asam_sound_ctx *ctx, void **ptr, size_t bytes);
int asam_dev_free(asam_sound_ctx *ctx, void *ptr);
int asam_dev_upload(asam_sound_ctx *ctx, void *dst, const void *src, size_t bytes);
int asam_dev_download(asam_sound_ctx *ctx, void *dst, const void *src, size_t bytes);
int asam_dev_memset(asam_sound_ctx *ctx, void *dst, int value, size_t bytes);
/* Device-side linear algebra used by Rosenbrock drivers (all on the context's stream):
 * y = a*x + b*y over n doubles; dot = sum(x*y) (FP64, deterministic order). */
int asam_dev_axpby(asam_sound_ctx *ctx, size_t n, double a, const double *x, double b,
                   double *y);
int asam_dev_dot(asam_sound_ctx *ctx, size_t n, const double *x, const double *y, double *dot);
/* y = V .* x (component-wise cell volume, the mass matrix of the stage system). */
int asam_dev_mass(asam_sound_ctx *ctx, double *y, const double *x);

/* Allocation-free form of asam_dev_dot.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code inside HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Needs no context: on the current HIP device it
 * enqueues sum(x*y) over n doubles on `stream` (a hipStream_t; NULL for the null stream)
 * and writes the result to the device double *dot, without synchronising; the sum has the
 * same bits as asam_dev_dot's.  Its partial sums live in a static device array of the
 * library (1024 doubles per device, held until the process ends); nothing is allocated.
 * Calls that overlap on one device (other streams or host threads) share that array and
 * race.  n at most ASAM_DOT_NOALLOC_MAX_N; returns ASAM_E_INVALID for larger n or a null
 * pointer, ASAM_E_HIP if the launch fails. */
#define ASAM_DOT_NOALLOC_MAX_N 4294967296LL /* 2^32, the longest vector a context holds */
int asam_dev_dot_noalloc(void *stream, size_t n, const double *x, const double *y,
                         double *dot);

/* ---- measurement hooks -------------------------------------------------- */
/* Average device time (ms, HIP events on the context's stream) of one launch of the
 * stencil apply (which=0), line preconditioner (which=1) or fused BiCGStab update
 * (which=2), in FP64 (fp32=0) or FP32 (fp32=1), or of the FP64 right-hand side (which=3)
 * or post-solve update (which=4, both of its kernels), over reps launches. */
int asam_sound_time_kernel(asam_sound_ctx *ctx, int which, int fp32, int reps, double *ms);
/* Average time (ms) of a 16-byte-vector streaming copy of `bytes` bytes (reads bytes, writes
 * bytes): the achievable-bandwidth and profiler-counter calibration kernel. */
int asam_sound_time_copy(asam_sound_ctx *ctx, size_t bytes, int reps, double *ms);
/* Kernel launch counts of the last solve, per kernel class (apply, line, blas, reduce). */
int asam_sound_last_counts(const asam_sound_ctx *ctx, int counts[8]);
/* Live per-kernel timing: when enabled, every launch of a solve is bracketed by HIP
 * events on the context's stream and its algorithmic bytes (each distinct array element
 * read or written once) are accounted.  Classes: 0 apply FP64, 1 apply FP32, 2 line FP64,
 * 3 line FP32, 4 BLAS-1 FP64, 5 BLAS-1 FP32, 6 reduction tails, 7 other.
 * asam_sound_profile_read returns the sums since the last read and resets them. */
int asam_sound_profile_enable(asam_sound_ctx *ctx, int on);
int asam_sound_profile_read(asam_sound_ctx *ctx, double ms[8], double bytes[8], int launches[8]);

#ifdef __cplusplus
}
#endif
#endif
