/*
 * asam_transport.h -- C-ABI of the MI355X (gfx950) transport stage solver.
 *
 * Drop-in for ASAM's transport stage solve, the second implicit system of every
 * Rosenbrock stage: BiCGStabChemSp / BiCGStabMetSp (reference
 * SRC/MODULE/SpMatrix4CellPar_Mod.f90:140-348) over the transport Jacobian
 * TYPE(JacSpMatrix4_T) of one block, with the reference's conventions kept here:
 *   - vectors are TYPE(Vector4Cell_T) of one block: ncomp components, each a haloed
 *     array c(ix0:ix1+1, iy0:iy1+1, iz0:iz1+1) of m = (nx+2)(ny+2)(nz+2) FP64 entries,
 *     x fastest, stored here as ncomp consecutive arrays; dots run over the interior
 *     cells only (reference SRC/MODULE/Vector4Cell_Mod.f90:1120-1132);
 *   - the transport matrices are TYPE(SpDiag) (reference SRC/CG/MatSpDiag_Mod.f90:7-13)
 *     in the layout of SpDiagAllocate (Vector4Cell_Mod.f90:1020-1063): Val(m, NumDiag)
 *     column-major, entry Val(i,d) = A(i - DiagPtr(d), i).  JacTMom and JacTPot have
 *     NumDiag = 7, DiagPtr = -(nx+2)(ny+2), -(nx+2), -1, 0, 1, nx+2, (nx+2)(ny+2),
 *     Diag = 4; JacFall has NumDiag = 2, DiagPtr = 0, (nx+2)(ny+2), Diag = 1;
 *   - the operator is MatVecChem_Vec4CellSp / MatVecMet_Vec4CellSp
 *     (SpMatrix4CellPar_Mod.f90:56-84): y = x + LU^-1 (T x - x) on the interior cells with
 *     the per-cell LU factors (JacSChemLU / JacSMetLU after GefaMatrix), T x on the halo;
 *   - the preconditioner is PSolve_Vec4CellSp -> PSolveSp (SpMatrix4CellPar_Mod.f90:86-102,
 *     Vector4Cell_Mod.f90:694-740): per component, 1 + IterG forward Gauss-Seidel sweeps
 *     (GaussSeidel_SpDiag / _SpDiagAB, MatSpDiag_Mod.f90:190-296) in the reference's
 *     lexicographic order, then the six halo faces set to zero.
 *
 * Error codes and the solve status are those of asam_sound.h.  Every function returns 0
 * or a negative ASAM_E_* code and never aborts the host.  A context is bound to the HIP
 * device current at creation and owns one HIP stream; calls on one context are not
 * re-entrant, calls on different contexts are independent.
 */
#ifndef ASAM_TRANSPORT_H
#define ASAM_TRANSPORT_H

#include <stddef.h>

#include "asam_sound.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Operator of one component (the branches of PSolveSp and MatVecMetSp,
 * Vector4Cell_Mod.f90:605-644, 694-740). */
#define ASAM_TR_MOM 0  /* JacTMom (components uPosL..wPosR) */
#define ASAM_TR_POT 1  /* JacTPot */
#define ASAM_TR_FALL 2 /* JacTPot + alpha * JacFall (ScaleMat2(ic) = alpha < 0) */

/* Matrix selector of asam_tr_set_matrix. */
#define ASAM_TR_JACTMOM 0
#define ASAM_TR_JACTPOT 1
#define ASAM_TR_JACFALL 2

#define ASAM_TR_MAX_COMP 16

typedef struct asam_tr_ctx asam_tr_ctx;

typedef struct {
  int nx, ny, nz; /* interior cells; nz <= 126 */
  int ncomp;      /* components of the vectors, 1..ASAM_TR_MAX_COMP */
  int iterg;      /* MatSpDiag_Mod IterG: sweeps after the first (reference 5), 0..32 */
  int reserved[3];
} asam_tr_grid;

typedef struct {
  int iterations; /* BiCGStab iterations begun (the half-step exit counts its iteration) */
  int status;     /* ASAM_CONVERGED, ASAM_NOT_CONVERGED or ASAM_BREAKDOWN */
  double rnorm0;  /* ||b - A x0|| over the interior cells */
  double rnorm;   /* final ||r|| */
  int psolves;    /* preconditioner applications */
  int matvecs;    /* operator applications, the initial residual included */
} asam_tr_info;

/* Launch shapes of the Gauss-Seidel sweep (results do not depend on them). */
typedef struct {
  int gs_rows;       /* y-rows per workgroup: 1..8 (one wavefront per row; two and at most
                        4 rows when nz > 62) */
  int gs_chunk;      /* steps whose coefficients a wavefront loads in one batch: 2 or 4 */
  int gs_spin_limit; /* polls of a boundary value before a block stops waiting for it (the
                        call then returns ASAM_E_HIP); 0: 4194304 */
  int reserved;
} asam_tr_tuning;

/* Creates the context and its device workspace (no per-call allocation). */
int asam_tr_create(asam_tr_ctx **ctx, const asam_tr_grid *grid);
int asam_tr_destroy(asam_tr_ctx *ctx);
const char *asam_tr_last_error(const asam_tr_ctx *ctx);

/* Values of JacTMom, JacTPot or JacFall (which = ASAM_TR_JAC*), as filled by JacMeteo /
 * JacScaleCompute (reference SRC/ADVECT/Function_Mod.f90:1508-1939): val is Val(m,
 * numdiag) column-major; numdiag, diag and diagptr must be the layout above.  The
 * Gauss-Seidel order needs every coupling to be between cells adjacent in the haloed
 * box, so entries that couple the last cell of an x-row (or y-plane) to the first of the
 * next are rejected (the reference's assembly never sets them). */
int asam_tr_set_matrix(asam_tr_ctx *ctx, int which, int numdiag, int diag,
                       const int *diagptr, const double *val);
/* Per-cell LU factors after GefaMatrix (Vector4Cell_Mod.f90:1404-1429): entry (ic,kc) of
 * interior cell q at lu[(ic*ncomp + kc)*nx*ny*nz + q], q x fastest.  NULL: identity. */
int asam_tr_set_lu(asam_tr_ctx *ctx, const double *lu);
/* Operator of every component in MatVec (op_mv) and PSolve (op_ps), ASAM_TR_*; alpha:
 * ScaleMat2 of each component (used by ASAM_TR_FALL).  BiCGStabChem: op_mv all
 * ASAM_TR_POT (MatVecChemSp); BiCGStabMet: op_mv = op_ps.  NULL arrays: all POT, 0. */
int asam_tr_set_components(asam_tr_ctx *ctx, const int *op_mv, const int *op_ps,
                           const double *alpha);

/* y = A x (MatVecChem/Met_Vec4CellSp); init != 0 takes JacTPot for every component, the
 * form BiCGStab uses for its initial residual (SpMatrix4CellPar_Mod.f90:164, 271).
 * Host buffers of ncomp*m doubles. */
int asam_tr_matvec(asam_tr_ctx *ctx, const double *x, double *y, int init);
/* y = P x (PSolve_Vec4CellSp); x and y may be the same buffer. */
int asam_tr_psolve(asam_tr_ctx *ctx, const double *x, double *y);
/* BiCGStabChem/Met_Vec4CellSp: x in/out (initial guess, solution); maxiter and tol in/out
 * with the reference's conventions: maxiter returns -Iter on the half-step exit, Iter on
 * the full-step exit, 2*Iter when not converged; tol returns ||r||; rho == 0 returns
 * maxiter = Iter, tol = 0 (status ASAM_BREAKDOWN). */
int asam_tr_solve(asam_tr_ctx *ctx, double *x, const double *b, int *maxiter, double *tol,
                  asam_tr_info *info);
/* Same on device buffers (any HIP allocation on the context's device). */
int asam_tr_matvec_device(asam_tr_ctx *ctx, const double *x, double *y, int init);
int asam_tr_psolve_device(asam_tr_ctx *ctx, const double *x, double *y);
int asam_tr_solve_device(asam_tr_ctx *ctx, double *x, const double *b, int *maxiter,
                         double *tol, asam_tr_info *info);

int asam_tr_get_tuning(const asam_tr_ctx *ctx, asam_tr_tuning *t);
int asam_tr_set_tuning(asam_tr_ctx *ctx, const asam_tr_tuning *t);
/* Device time of the kernels on the context's own buffers, averaged over reps launches
 * (HIP events on the context's stream): which 0 = one Gauss-Seidel sweep over all
 * components, 1 = the operator (MatVec), 2 = one whole PSolve.  No counterpart in the
 * reference (measurement). */
int asam_tr_time_kernel(asam_tr_ctx *ctx, int which, int reps, double *ms);

/* Test hook, no counterpart in the reference: with hold != 0 no block publishes its last
 * row, so every later block waits for its boundary values until gs_spin_limit polls and
 * the sweep returns ASAM_E_HIP, every kernel still finishing normally.  For tests of the
 * bounded wait (with a small gs_spin_limit); hold = 0 restores normal sweeps. */
int asam_tr_test_hold_boundary(asam_tr_ctx *ctx, int hold);

/* Device buffers on the context's device (for the *_device entry points). */
int asam_tr_dev_alloc(asam_tr_ctx *ctx, void **ptr, size_t bytes);
int asam_tr_dev_free(asam_tr_ctx *ctx, void *ptr);
int asam_tr_dev_upload(asam_tr_ctx *ctx, void *dst, const void *src, size_t bytes);
int asam_tr_dev_download(asam_tr_ctx *ctx, void *dst, const void *src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
