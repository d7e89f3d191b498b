/*
 * transport.c -- CPU restatement (FP64) of ASAM's transport stage solve.
 *
 * TEST INFRASTRUCTURE ONLY (see transport.h).  The loops keep the reference's own order of
 * operations, including the column-oriented sweeps of the SpDiag routines, so that the
 * arithmetic of every entry is the reference's:
 *   - SpAVec_SpDiag / SpABVec_SpDiag        CG/MatSpDiag_Mod.f90:95-143
 *   - GaussSeidel_SpDiag / _SpDiagAB         CG/MatSpDiag_Mod.f90:190-296 (IterG :14)
 *   - PSolveMatrix, GefaMatrix              MODULE/Vector4Cell_Mod.f90:742-776, 1404-1429
 *   - MatVecChemSp / MatVecMetSp / PSolveSp MODULE/Vector4Cell_Mod.f90:605-740
 *   - Dot_VectorCell, Axpy/Xpay/Xpy/Copy    MODULE/Vector4Cell_Mod.f90:1120-1340
 *   - MatVecChem/Met_Vec4CellSp, PSolve_Vec4CellSp, BiCGStabChem/Met_Vec4CellSp
 *                                           MODULE/SpMatrix4CellPar_Mod.f90:56-348
 * Built with ISO C (no floating-point contraction), so products and sums round
 * separately, as in the reference.
 */
#include "transport.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Val(i, d), 1-based as in the reference. */
#define VAL(A, i, d) ((A)->val[(long)((d)-1) * (A)->n + ((i)-1)])

void orc_tr_spavec(const orc_spdiag *A, const orc_spdiag *B, double alpha, double *y,
                   const double *x) {
  const long n = A->n;
  for (long i = 0; i < n; ++i) y[i] = 0.0;
  for (int d = 1; d <= A->numdiag; ++d) {
    const long s = A->ptr[d - 1];
    const long lo = 1 + (s > 0 ? s : 0), hi = n + (s < 0 ? s : 0);
#pragma omp parallel for schedule(static)
    for (long i = lo; i <= hi; ++i) y[i - s - 1] += VAL(A, i, d) * x[i - 1];
  }
  if (B) {
    for (int d = 1; d <= B->numdiag; ++d) {
      const long s = B->ptr[d - 1];
      const long lo = 1 + (s > 0 ? s : 0), hi = B->n + (s < 0 ? s : 0);
#pragma omp parallel for schedule(static)
      for (long i = lo; i <= hi; ++i) y[i - s - 1] += (VAL(B, i, d) * alpha) * x[i - 1];
    }
  }
}

static double gs_diag(const orc_spdiag *A, const orc_spdiag *C, double alpha, long i) {
  if (!C) return VAL(A, i, A->diag);
  return VAL(A, i, A->diag) + VAL(C, i, C->diag) * alpha;
}

/* Forward substitution (L+D) of one sweep, column oriented (:205-212, :282-293). */
static void gs_forward(const orc_spdiag *A, const orc_spdiag *C, double alpha, double *T) {
  const long n = A->n;
  for (long i = 1; i <= n; ++i) {
    T[i] = T[i] / gs_diag(A, C, alpha, i);
    for (int d = 1; d <= A->diag - 1; ++d) T[i - A->ptr[d - 1]] -= VAL(A, i, d) * T[i];
    if (C)
      for (int d = 1; d <= C->diag - 1; ++d)
        T[i - C->ptr[d - 1]] -= (VAL(C, i, d) * alpha) * T[i];
  }
}

void orc_tr_gauss_seidel(const orc_spdiag *A, const orc_spdiag *C, double alpha, int iterg,
                         double *x, const double *b) {
  const long n = A->n;
  long pad = -(long)A->ptr[0];  /* Temp(1+DiagPtr(1) : n-DiagPtr(1)) */
  if (pad < 0) pad = -pad;
  double *buf = calloc((size_t)(n + 2 * pad + 1), sizeof(double));
  double *T = buf + pad; /* T[i], i in 1-pad .. n+pad */
  for (long i = 1; i <= n; ++i) T[i] = b[i - 1];
  gs_forward(A, C, alpha, T);
  for (int it = 1; it <= iterg; ++it) {
    /* T := b - U T (:217-223, :267-279): row i-DiagPtr receives the old T(i) */
    for (long i = 1; i <= n; ++i) {
      for (int d = A->diag + 1; d <= A->numdiag; ++d) T[i - A->ptr[d - 1]] -= VAL(A, i, d) * T[i];
      if (C)
        for (int d = C->diag + 1; d <= C->numdiag; ++d)
          T[i - C->ptr[d - 1]] -= (VAL(C, i, d) * alpha) * T[i];
      T[i] = b[i - 1];
    }
    gs_forward(A, C, alpha, T);
  }
  memcpy(x, T + 1, (size_t)n * sizeof(double));
  free(buf);
}

void orc_tr_lu_solve(int nx, int ny, int nz, int nc, const double *lu, double *v) {
  const long X = nx + 2, Y = ny + 2, Z = nz + 2, m = X * Y * Z;
  const long ncell = (long)nx * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const long r = i + X * (j + Y * (long)k);
        const long q = (i - 1) + (long)nx * ((j - 1) + (long)ny * (k - 1));
#define LU(ic, kc) lu[((long)(ic)*nc + (kc)) * ncell + q]
#define B(ic) v[(long)(ic)*m + r]
        for (int kc = 0; kc < nc; ++kc)
          for (int ic = kc + 1; ic < nc; ++ic) B(ic) = B(ic) - LU(ic, kc) * B(kc);
        for (int kc = nc - 1; kc >= 0; --kc) {
          B(kc) = B(kc) / LU(kc, kc);
          for (int ic = 0; ic < kc; ++ic) B(ic) = B(ic) - LU(ic, kc) * B(kc);
        }
#undef LU
#undef B
      }
  (void)Z;
}

void orc_tr_gefa(long ncell, int nc, double fac, double *lu) {
#pragma omp parallel for schedule(static)
  for (long q = 0; q < ncell; ++q) {
#define LU(ic, kc) lu[((long)(ic)*nc + (kc)) * ncell + q]
    for (int ic = 0; ic < nc; ++ic) {
      for (int jc = 0; jc < nc; ++jc) LU(ic, jc) = -(fac * LU(ic, jc));
      LU(ic, ic) = 1.0 + LU(ic, ic);
    }
    for (int kc = 0; kc < nc; ++kc)
      for (int ic = kc + 1; ic < nc; ++ic) {
        LU(ic, kc) = LU(ic, kc) / LU(kc, kc);
        for (int jc = kc + 1; jc < nc; ++jc) LU(ic, jc) = LU(ic, jc) - LU(ic, kc) * LU(kc, jc);
      }
#undef LU
  }
}

struct orc_tr {
  int nx, ny, nz, nc, iterg;
  long m;
  int ptr7[7], ptr2[2];
  orc_spdiag tmom, tpot, fall;
  int has_fall;
  const double *lu;
  int *op_mv, *op_ps;
  double *alpha;
};

orc_tr *orc_tr_create(int nx, int ny, int nz, int nc, int iterg, const double *tmom,
                      const double *tpot, const double *fall, const double *lu,
                      const int *op_mv, const int *op_ps, const double *alpha) {
  orc_tr *T = calloc(1, sizeof(orc_tr));
  T->nx = nx, T->ny = ny, T->nz = nz, T->nc = nc, T->iterg = iterg;
  const long X = nx + 2, XY = X * (ny + 2);
  T->m = XY * (nz + 2);
  const int p7[7] = {(int)-XY, (int)-X, -1, 0, 1, (int)X, (int)XY};
  memcpy(T->ptr7, p7, sizeof p7);
  T->ptr2[0] = 0, T->ptr2[1] = (int)XY;
  T->tmom = (orc_spdiag){(int)T->m, 7, 4, T->ptr7, tmom};
  T->tpot = (orc_spdiag){(int)T->m, 7, 4, T->ptr7, tpot};
  T->fall = (orc_spdiag){(int)T->m, 2, 1, T->ptr2, fall};
  T->has_fall = fall != NULL;
  T->lu = lu;
  T->op_mv = malloc(sizeof(int) * nc);
  T->op_ps = malloc(sizeof(int) * nc);
  T->alpha = malloc(sizeof(double) * nc);
  for (int c = 0; c < nc; ++c) {
    T->op_mv[c] = op_mv ? op_mv[c] : ORC_TR_POT;
    T->op_ps[c] = op_ps ? op_ps[c] : ORC_TR_POT;
    T->alpha[c] = alpha ? alpha[c] : 0.0;
  }
  return T;
}

void orc_tr_free(orc_tr *T) {
  if (!T) return;
  free(T->op_mv);
  free(T->op_ps);
  free(T->alpha);
  free(T);
}

long orc_tr_m(const orc_tr *T) { return T->m; }

static const orc_spdiag *op_mat(const orc_tr *T, int op) {
  return op == ORC_TR_MOM ? &T->tmom : &T->tpot;
}

void orc_tr_matvec(const orc_tr *T, double *y, const double *x, int init) {
  const long m = T->m, M = m * T->nc;
  for (int c = 0; c < T->nc; ++c) {
    const int op = init ? ORC_TR_POT : T->op_mv[c];
    const orc_spdiag *B = (op == ORC_TR_FALL && T->has_fall) ? &T->fall : NULL;
    orc_tr_spavec(op_mat(T, op), B, T->alpha[c], y + c * m, x + c * m);
  }
#pragma omp parallel for schedule(static)
  for (long e = 0; e < M; ++e) y[e] = -1.0 * x[e] + y[e]; /* Axpy(-One, x, Ax) */
  if (T->lu) orc_tr_lu_solve(T->nx, T->ny, T->nz, T->nc, T->lu, y);
#pragma omp parallel for schedule(static)
  for (long e = 0; e < M; ++e) y[e] = x[e] + y[e]; /* Xpy(x, Ax) */
}

static void zero_halo(const orc_tr *T, double *v) {
  const long X = T->nx + 2, Y = T->ny + 2, Z = T->nz + 2;
  for (long k = 0; k < Z; ++k)
    for (long j = 0; j < Y; ++j)
      for (long i = 0; i < X; ++i)
        if (i == 0 || i == X - 1 || j == 0 || j == Y - 1 || k == 0 || k == Z - 1)
          v[i + X * (j + Y * k)] = 0.0;
}

void orc_tr_psolve(const orc_tr *T, double *y, const double *x) {
  const long m = T->m;
  if (y != x) memcpy(y, x, sizeof(double) * m * T->nc); /* Copy(x, Px) */
  for (int c = 0; c < T->nc; ++c) {
    const int op = T->op_ps[c];
    const orc_spdiag *C = (op == ORC_TR_FALL && T->has_fall) ? &T->fall : NULL;
    orc_tr_gauss_seidel(op_mat(T, op), C, T->alpha[c], T->iterg, y + c * m, y + c * m);
    zero_halo(T, y + c * m);
  }
}

double orc_tr_dot(const orc_tr *T, const double *x, const double *y) {
  const long X = T->nx + 2, Y = T->ny + 2, m = T->m;
  double total = 0.0;
  for (int c = 0; c < T->nc; ++c) {
    double s = 0.0; /* SUM over the interior section, column-major order */
    for (long k = 1; k <= T->nz; ++k)
      for (long j = 1; j <= T->ny; ++j)
        for (long i = 1; i <= T->nx; ++i) {
          const long r = c * m + i + X * (j + Y * k);
          s += x[r] * y[r];
        }
    total = total + s;
  }
  return total;
}

/* Axpy_VectorCell: skipped when alpha == 0 (:1243). */
static void axpy(long M, double a, const double *x, double *y) {
  if (a == 0.0) return;
#pragma omp parallel for schedule(static)
  for (long e = 0; e < M; ++e) y[e] = a * x[e] + y[e];
}
/* Xpay_VectorCell: y = a*y + x (:1316-1326). */
static void xpay(long M, const double *x, double a, double *y) {
#pragma omp parallel for schedule(static)
  for (long e = 0; e < M; ++e) y[e] = a * y[e] + x[e];
}

int orc_tr_bicgstab(const orc_tr *T, double *x, const double *b, int *maxiter, double *tol) {
  const long M = T->m * T->nc;
  const size_t bytes = sizeof(double) * (size_t)M;
  double *r = malloc(bytes), *rt = malloc(bytes), *rh = malloc(bytes), *p = malloc(bytes);
  double *ph = malloc(bytes), *t = malloc(bytes), *v = malloc(bytes);
  int status = 0;
  const int MaxIter = *maxiter;
  const double Tol = *tol;
  orc_tr_matvec(T, r, x, 1); /* MatVecChem(r, A, x) (:164, :271) */
  xpay(M, b, -1.0, r);       /* r = b - r */
  memcpy(rt, r, bytes);
  const double rnorm0 = sqrt(orc_tr_dot(T, r, r));
  double rho = 0.0, rho_old, alpha = 0.0, omega = 0.0, beta, rnorm;
  for (int it = 1; it <= MaxIter; ++it) {
    rho_old = rho;
    rho = orc_tr_dot(T, rt, r);
    if (rho == 0.0) {
      *maxiter = it;
      *tol = 0.0;
      status = 2;
      break;
    }
    if (it == 1) {
      memcpy(p, r, bytes);
    } else {
      beta = (rho / rho_old) * (alpha / omega);
      axpy(M, -omega, v, p);
      xpay(M, r, beta, p);
    }
    orc_tr_psolve(T, ph, p);
    orc_tr_matvec(T, v, ph, 0);
    alpha = rho / orc_tr_dot(T, rt, v);
    axpy(M, -alpha, v, r);
    rnorm = sqrt(orc_tr_dot(T, r, r));
    if (rnorm <= Tol * rnorm0) {
      axpy(M, alpha, ph, x);
      *maxiter = -it;
      *tol = rnorm;
      break;
    }
    orc_tr_psolve(T, rh, r);
    orc_tr_matvec(T, t, rh, 0);
    omega = orc_tr_dot(T, t, r) / orc_tr_dot(T, t, t);
    axpy(M, alpha, ph, x);
    axpy(M, omega, rh, x);
    axpy(M, -omega, t, r);
    rnorm = sqrt(orc_tr_dot(T, r, r));
    if (rnorm <= Tol * rnorm0) {
      *maxiter = it;
      *tol = rnorm;
      break;
    }
    if (it == MaxIter) {
      *maxiter = 2 * it;
      *tol = rnorm;
      break;
    }
  }
  free(r), free(rt), free(rh), free(p), free(ph), free(t), free(v);
  return status;
}
