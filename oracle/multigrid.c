/*
 * multigrid.c -- CPU restatement (FP64) of the reference's own preconditioner of the sound
 * stage solve, PressurePre: a semicoarsened multigrid whose smoother is a symmetric
 * lexicographic block Gauss-Seidel sweep over y-planes; each plane is solved inexactly by a
 * 2D multigrid whose smoother is a symmetric x-line Gauss-Seidel sweep with exact z-line
 * (Thomas) solves.
 *
 * TEST INFRASTRUCTURE ONLY: gives the iteration counts and iterates of the reference's
 * default preconditioner to compare the vertical-line preconditioner with.  The product
 * path never links or calls this file.
 *
 * Restated (reference paths relative to SRC/):
 *   - PressurePre MATRIX/PreCond_Mod.f90:599-622 -> VCycleCG MULTI/JacAccGrav_Mod.f90:4468-4517
 *     (MultiTriT: x = 0, Solve_3D; MultiTriTB: x = 0, one Multigrid_3D); PSolve1
 *     (PreCond_Mod.f90:1555-1571) is -PressurePre for one closed block (the first
 *     PressurePre only feeds interface terms, b1pBTx2 then sets b%p = -x%p).
 *   - hierarchy: LaplTDualCompute / LaplTFullCompute / LaplTBCompute
 *     (JacAccGrav_Mod.f90:902-1484): 3D level L3 = 0..RefLevel-1 coarsens all three
 *     directions by 2^L3; each of its planes carries 2D levels L2 = L3+1..RefLevel-1 that
 *     coarsen z and x by 2^L2 and keep that level's y resolution.  Coarse data by
 *     FaceCoarse / CFaceCoarse / VolCoarse / CVolCoarse / IncrCoarse / CIncrCoarse
 *     (:607-891) from the finest arrays, at level 0 too (shift 0).
 *   - coefficients: ComputeDual_1D/2D/3D, ComputeFull_1D/2D/3D (MULTI/TriDiagT_Mod.f90
 *     :32-90, 253-399, 693-900) and Compute_1D/2D/3D of MultiTriTB (TriDiagTB_Mod.f90:32-93,
 *     261-343, 663-780), with Allocate_1D's 1e-40 diagonal seed; the 3D couplings of a
 *     level with its own y couplings omit DUU3 as the reference's first branch does.
 *   - cycles: GaussSeidel_2D/3D (TriDiagT_Mod.f90:430-474, 951-1007, with the extra
 *     top-plane solve :998-1001; TriDiagTB_Mod.f90:374-418, 818-874), Solve_2D/3D
 *     (MultIter2/3 cycles), Multigrid_2D/3D (GaussIter grows by GaussIter2/3 per level on
 *     the way down, the coarsest level smoothed twice), MatVecT/TB_1D/2D/3D, RestrictT
 *     (summation) / ProlongateT (piecewise constant) and the TB forms RestrictTB /
 *     ProlongateTB with their summation orders, Solve_1D (Thomas, 2x2 blocks by Cramer's
 *     rule, CG/Matrix_Mod.f90:34-94).
 *   - axis roles xOrder = 2, yOrder = 3, zOrder = 1 (MODULE/Control_Mod.f90:601-603): the
 *     reversed index order is n1 = z, n2 = x, n3 = y; coarsening factor 2 in every direction
 *     (xCoarse = yCoarse = zCoarse = 2, :604-606); closed walls (SetFactorN,
 *     Reverse_Mod.f90:215-297, all boundary factors zero, so wall faces add nothing); the
 *     Cartesian metric makes every coarse metric factor exactly 1 (CIncrCoarse averages 1
 *     with the same sums IncrCoarse forms).
 *
 * Vectors inside: component fastest, then i1 (z), i2 (x), i3 (y).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EPS 1.0e-30

typedef struct { /* one 2D level of every plane of a 3D level */
  int m1, m2, m3;
  double *d, *u1, *l1, *u2, *l2;
  /* work of one plane's V-cycle at this level (planes are solved one at a time) */
  double *res, *rc, *xc, *t, *w;
} lev2;

typedef struct { /* one 3D level */
  int m1, m2, m3, nq;
  lev2 *q;         /* q[0]: the planes of this level; q[1..]: their coarse 2D levels */
  double *u3, *l3; /* y couplings */
  double *res, *rc, *xc, *t;
} lev3;

struct orc_mg {
  orc_grid g;
  orc_mg_control ctl;
  int np, n1, n2, n3, R;
  lev3 *L;
  double *xr, *br; /* reversed-order copies for pressurepre */
};

/* ------------------------------------------------------------------ small 2x2 helpers */
/* CG/Matrix_Mod.f90:34-94, column-major [11,21,12,22] */
static void mul22(double *r, const double *b, const double *c) {
  double t0 = b[0] * c[0] + b[2] * c[1];
  double t2 = b[0] * c[2] + b[2] * c[3];
  double t1 = b[1] * c[0] + b[3] * c[1];
  double t3 = b[1] * c[2] + b[3] * c[3];
  r[0] = t0; r[1] = t1; r[2] = t2; r[3] = t3;
}
static void mul2(double *r, const double *b, const double *c) {
  double t0 = b[0] * c[0] + b[2] * c[1];
  double t1 = b[1] * c[0] + b[3] * c[1];
  r[0] = t0; r[1] = t1;
}
static void div22(double *r, const double *b, const double *c) {
  double det = c[0] * c[3] - c[2] * c[1];
  double t0 = (c[3] * b[0] - c[2] * b[1]) / det;
  double t2 = (c[3] * b[2] - c[2] * b[3]) / det;
  double t1 = (-c[1] * b[0] + c[0] * b[1]) / det;
  double t3 = (-c[1] * b[2] + c[0] * b[3]) / det;
  r[0] = t0; r[1] = t1; r[2] = t2; r[3] = t3;
}
static void div2(double *r, const double *b, const double *c) {
  double det = c[0] * c[3] - c[2] * c[1];
  double t0 = (c[3] * b[0] - c[2] * b[1]) / det;
  double t1 = (-c[1] * b[0] + c[0] * b[1]) / det;
  r[0] = t0; r[1] = t1;
}

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }
static int cdiv(int n, int s) { return (n + s - 1) / s; }

/* ---------------------------------------------------------------- coarse grid data */
typedef struct {
  int m1, m2, m3;
  double *F1, *F2, *F3, *T1, *T2, *T3, *U1, *U2, *U3, *G1, *G2, *G3, *V, *DUT;
  double *dd1, *dd2, *dd3;
} cgrid;

/* face arrays: F1 (m1+1, m2, m3), F2 (m1, m2+1, m3), F3 (m1, m2, m3+1), i1 fastest */
#define I1(a, b, c, M1, M2) ((a) + (long)((M1) + 1) * ((b) + (long)(M2) * (c)))
#define I2(a, b, c, M1, M2) ((a) + (long)(M1) * ((b) + (long)((M2) + 1) * (c)))
#define I3(a, b, c, M1, M2) ((a) + (long)(M1) * ((b) + (long)(M2) * (c)))

static void cgrid_free(cgrid *C) {
  free(C->F1); free(C->F2); free(C->F3); free(C->T1); free(C->T2); free(C->T3);
  free(C->U1); free(C->U2); free(C->U3); free(C->G1); free(C->G2); free(C->G3);
  free(C->V); free(C->DUT); free(C->dd1); free(C->dd2); free(C->dd3);
  memset(C, 0, sizeof *C);
}

/* The coarse grid of shifts s1, s2, s3 (each 2^L) from the finest data f (which is the
 * shift-1 grid): FaceCoarse / CFaceCoarse / VolCoarse / CVolCoarse / IncrCoarse, their
 * loop orders kept, so every coarse sum accumulates in the reference's order.  TB weights
 * DUU by FG (CFaceCoarse(DUU, F?G), JacAccGrav_Mod.f90:1409-1411), T by F. */
static void coarsen(const cgrid *f, int tb, int s1, int s2, int s3, cgrid *C) {
  const int n1 = f->m1, n2 = f->m2, n3 = f->m3;
  const int m1 = cdiv(n1, s1), m2 = cdiv(n2, s2), m3 = cdiv(n3, s3);
  C->m1 = m1; C->m2 = m2; C->m3 = m3;
  const long nf1 = (long)(m1 + 1) * m2 * m3, nf2 = (long)m1 * (m2 + 1) * m3,
             nf3 = (long)m1 * m2 * (m3 + 1), nc = (long)m1 * m2 * m3;
  C->F1 = zalloc(nf1); C->T1 = zalloc(nf1); C->U1 = zalloc(nf1); C->G1 = zalloc(nf1);
  C->F2 = zalloc(nf2); C->T2 = zalloc(nf2); C->U2 = zalloc(nf2); C->G2 = zalloc(nf2);
  C->F3 = zalloc(nf3); C->T3 = zalloc(nf3); C->U3 = zalloc(nf3); C->G3 = zalloc(nf3);
  C->V = zalloc(nc); C->DUT = zalloc(nc);
  C->dd1 = zalloc(m1); C->dd2 = zalloc(m2); C->dd3 = zalloc(m3);
  /* weights of DUU: F (T) or FG (TB) */
  const double *W1 = tb ? f->G1 : f->F1, *W2 = tb ? f->G2 : f->F2, *W3 = tb ? f->G3 : f->F3;
  double *WC1 = zalloc(nf1), *WC2 = zalloc(nf2), *WC3 = zalloc(nf3);
  /* direction 1 (z faces): DO iC; i = MIN(s1*iC, n1); DO j; DO k */
  for (int iC = 0; iC <= m1; ++iC) {
    const int i = s1 * iC < n1 ? s1 * iC : n1;
    for (int j = 0; j < n2; ++j)
      for (int k = 0; k < n3; ++k) {
        const long fc = I1(iC, j / s2, k / s3, m1, m2), ff = I1(i, j, k, n1, n2);
        C->F1[fc] += f->F1[ff];
        C->T1[fc] += f->T1[ff] * f->F1[ff];
        C->U1[fc] += f->U1[ff] * W1[ff];
        WC1[fc] += W1[ff];
        if (tb) C->G1[fc] += f->G1[ff];
      }
  }
  /* direction 2 (x faces): DO i; DO jC; j = MIN(s2*jC, n2); DO k */
  for (int i = 0; i < n1; ++i)
    for (int jC = 0; jC <= m2; ++jC) {
      const int j = s2 * jC < n2 ? s2 * jC : n2;
      for (int k = 0; k < n3; ++k) {
        const long fc = I2(i / s1, jC, k / s3, m1, m2), ff = I2(i, j, k, n1, n2);
        C->F2[fc] += f->F2[ff];
        C->T2[fc] += f->T2[ff] * f->F2[ff];
        C->U2[fc] += f->U2[ff] * W2[ff];
        WC2[fc] += W2[ff];
        if (tb) C->G2[fc] += f->G2[ff];
      }
    }
  /* direction 3 (y faces): DO i; DO j; DO kC; k = MIN(s3*kC, n3) */
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n2; ++j)
      for (int kC = 0; kC <= m3; ++kC) {
        const int k = s3 * kC < n3 ? s3 * kC : n3;
        const long fc = I3(i / s1, j / s2, kC, m1, m2), ff = I3(i, j, k, n1, n2);
        C->F3[fc] += f->F3[ff];
        C->T3[fc] += f->T3[ff] * f->F3[ff];
        C->U3[fc] += f->U3[ff] * W3[ff];
        WC3[fc] += W3[ff];
        if (tb) C->G3[fc] += f->G3[ff];
      }
  /* CFaceCoarse: C = sum(C*F)/(sum(F)+Eps) */
  for (long q = 0; q < nf1; ++q) { C->T1[q] = C->T1[q] / (C->F1[q] + EPS); C->U1[q] = C->U1[q] / (WC1[q] + EPS); }
  for (long q = 0; q < nf2; ++q) { C->T2[q] = C->T2[q] / (C->F2[q] + EPS); C->U2[q] = C->U2[q] / (WC2[q] + EPS); }
  for (long q = 0; q < nf3; ++q) { C->T3[q] = C->T3[q] / (C->F3[q] + EPS); C->U3[q] = C->U3[q] / (WC3[q] + EPS); }
  free(WC1); free(WC2); free(WC3);
  /* VolCoarse, CVolCoarse: DO i; DO j; DO k */
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n2; ++j)
      for (int k = 0; k < n3; ++k) {
        const long c = I3(i / s1, j / s2, k / s3, m1, m2), cf = I3(i, j, k, n1, n2);
        C->V[c] += f->V[cf];
        C->DUT[c] += f->DUT[cf] * f->V[cf];
      }
  for (long c = 0; c < nc; ++c) C->DUT[c] = C->DUT[c] / (C->V[c] + EPS);
  /* IncrCoarse */
  for (int i = 0; i < n1; ++i) C->dd1[i / s1] += f->dd1[i];
  for (int j = 0; j < n2; ++j) C->dd2[j / s2] += f->dd2[j];
  for (int k = 0; k < n3; ++k) C->dd3[k / s3] += f->dd3[k];
}

/* -------------------------------------------------------------- coefficient compute */
/* Coefficients of the faces of plane k (i3 = k) of grid C into 2D level L (plane k) and,
 * when u3 != NULL, the y couplings between planes; the 3D sums follow Compute_3D: the
 * planes are computed one after another, the y-face terms of face k+1/2 added to plane k
 * before plane k+1 is computed and to plane k+1 after. */
typedef struct { const orc_grid *g; int np; } cctx;

static void line1d(const cctx *X, const cgrid *C, lev2 *L, int j, int k) {
  const int m1 = L->m1, m2 = L->m2, np = X->np;
  const double bd2 = X->g->beta0_dt * X->g->beta0_dt, G = X->g->grav, fa = X->g->fac_anela;
  const int full = X->g->grad_full;
  const long base = (long)m1 * (j + (long)m2 * k);
  double *d = L->d + base * np * np, *u = L->u1 + base * np * np, *l = L->l1 + base * np * np;
  const double *V = C->V + base, *DUT = C->DUT + base;
  for (int i = 0; i < m1; ++i) {  /* Allocate_1D */
    if (np == 1) d[i] = 1.0e-40;
    else { d[4 * i] = 1.0e-40; d[4 * i + 1] = 0.0; d[4 * i + 2] = 0.0; d[4 * i + 3] = 1.0e-40; }
    if (np == 1) { u[i] = 0.0; l[i] = 0.0; } else for (int q = 0; q < 4; ++q) { u[4 * i + q] = 0.0; l[4 * i + q] = 0.0; }
  }
  for (int i = 0; i + 1 < m1; ++i) {
    const long ff = I1(i + 1, j, k, m1, m2);
    const double F = C->F1[ff], T = C->T1[ff], U = C->U1[ff];
    if (np == 1) {
      double a, b;  /* the face term with DUT of the upper / lower cell */
      if (full) {
        const double den = (C->dd1[i] + C->dd1[i + 1]) * 1.0 * 1.0;
        a = bd2 * 2.0 * F * T * U * DUT[i + 1] / den;
        b = bd2 * 2.0 * F * T * U * DUT[i] / den;
      } else {
        const double den = V[i] + V[i + 1] + EPS;
        a = bd2 * 2.0 * F * T * U * F * DUT[i + 1] / den;
        b = bd2 * 2.0 * F * T * U * F * DUT[i] / den;
      }
      u[i] = -a;
      l[i] = -b;
      d[i] = d[i] + b;
      d[i + 1] = d[i + 1] + a;
    } else {
      const double FG = C->G1[ff], den = V[i] + V[i + 1] + EPS;
      double *uu = u + 4 * i, *ll = l + 4 * i, *d0 = d + 4 * i, *d1 = d + 4 * (i + 1);
      uu[0] = -(bd2 * 2.0 * F * T * U * FG * DUT[i + 1] / den);
      uu[1] = -(bd2 * 2.0 * F * U * FG * DUT[i + 1] / den);
      ll[0] = -(bd2 * 2.0 * F * T * U * FG * DUT[i] / den);
      ll[1] = -(bd2 * 2.0 * F * U * FG * DUT[i] / den);
      uu[2] = uu[2] - G * bd2 * F * T * V[i + 1] / den;
      uu[3] = uu[3] - G * bd2 * F * V[i + 1] / den;
      ll[2] = ll[2] + G * bd2 * F * T * V[i] / den;
      ll[3] = ll[3] + G * bd2 * F * V[i] / den;
      d0[0] = d0[0] + bd2 * 2.0 * F * T * U * FG * DUT[i] / den;
      d0[1] = d0[1] + bd2 * 2.0 * F * U * FG * DUT[i] / den;
      d1[0] = d1[0] + bd2 * 2.0 * F * T * U * FG * DUT[i + 1] / den;
      d1[1] = d1[1] + bd2 * 2.0 * F * U * FG * DUT[i + 1] / den;
      d0[2] = d0[2] - G * bd2 * F * T * V[i] / den;
      d0[3] = d0[3] - G * bd2 * F * V[i] / den;
      d1[2] = d1[2] + G * bd2 * F * T * V[i + 1] / den;
      d1[3] = d1[3] + G * bd2 * F * V[i + 1] / den;
    }
  }
  for (int i = 0; i < m1; ++i) {
    if (np == 1) d[i] = d[i] + fa * V[i];
    else { d[4 * i] = d[4 * i] + fa * V[i]; d[4 * i + 3] = d[4 * i + 3] + V[i]; }
  }
}

/* x-face (j+1/2) term of cell (i, j|j+1) of plane k: add to d of line j (lower = 1) or
 * j+1 (lower = 0), set u2/l2 (Compute_2D / ComputeDual_2D / ComputeFull_2D). */
static void xface(const cctx *X, const cgrid *C, lev2 *L, int j, int k, int lower, int set) {
  const int m1 = L->m1, m2 = L->m2, np = X->np;
  const double bd2 = X->g->beta0_dt * X->g->beta0_dt;
  const int full = X->g->grad_full;
  for (int i = 0; i < m1; ++i) {
    const long ff = I2(i, j + 1, k, m1, m2);
    const long ca = I3(i, j, k, m1, m2), cb = I3(i, j + 1, k, m1, m2);
    const double F = C->F2[ff], T = C->T2[ff], U = C->U2[ff];
    const long cell = lower ? ca : cb;
    if (np == 1) {
      double a, b;
      if (full) {
        const double den = (C->dd2[j] + C->dd2[j + 1]) * 1.0 * 1.0;
        a = bd2 * 2.0 * F * T * U * C->DUT[cb] / den;
        b = bd2 * 2.0 * F * T * U * C->DUT[ca] / den;
      } else {
        const double den = C->V[ca] + C->V[cb] + EPS;
        a = bd2 * 2.0 * F * T * U * F * C->DUT[cb] / den;
        b = bd2 * 2.0 * F * T * U * F * C->DUT[ca] / den;
      }
      if (set) { L->u2[ca] = -a; L->l2[ca] = -b; }
      L->d[cell] = L->d[cell] + (lower ? b : a);
    } else {
      const double FG = C->G2[ff], den = C->V[ca] + C->V[cb] + EPS;
      const double D = lower ? C->DUT[ca] : C->DUT[cb];
      if (set) {
        L->u2[2 * ca] = -(bd2 * 2.0 * F * T * U * FG * C->DUT[cb] / den);
        L->u2[2 * ca + 1] = -(bd2 * 2.0 * F * U * FG * C->DUT[cb] / den);
        L->l2[2 * ca] = -(bd2 * 2.0 * F * T * U * FG * C->DUT[ca] / den);
        L->l2[2 * ca + 1] = -(bd2 * 2.0 * F * U * FG * C->DUT[ca] / den);
      }
      L->d[4 * cell] = L->d[4 * cell] + bd2 * 2.0 * F * T * U * FG * D / den;
      L->d[4 * cell + 1] = L->d[4 * cell + 1] + bd2 * 2.0 * F * U * FG * D / den;
    }
  }
}

static void plane2d(const cctx *X, const cgrid *C, lev2 *L, int k) {
  const int m2 = L->m2;
  line1d(X, C, L, 0, k);
  for (int j = 0; j + 1 < m2; ++j) {
    xface(X, C, L, j, k, 1, 1);
    line1d(X, C, L, j + 1, k);
    xface(X, C, L, j, k, 0, 0);
  }
}

/* y-face (k+1/2) term into plane k (lower) or k+1; with_ul: the level's own y couplings
 * (first branch of Compute_3D, without DUU3) */
static void yface(const cctx *X, const cgrid *C, lev2 *L, double *u3, double *l3, int k,
                  int lower) {
  const int m1 = L->m1, m2 = L->m2, np = X->np;
  const double bd2 = X->g->beta0_dt * X->g->beta0_dt;
  const int full = X->g->grad_full;
  const int with_ul = u3 != NULL;
  for (int j = 0; j < m2; ++j)
    for (int i = 0; i < m1; ++i) {
      const long ff = I3(i, j, k + 1, m1, m2);
      const long ca = I3(i, j, k, m1, m2), cb = I3(i, j, k + 1, m1, m2);
      const double F = C->F3[ff], T = C->T3[ff], U = C->U3[ff];
      const long cell = lower ? ca : cb;
      if (np == 1) {
        double a, b;
        if (full) {
          const double den = (C->dd3[k] + C->dd3[k + 1]) * 1.0 * 1.0;
          if (with_ul) {
            a = bd2 * 2.0 * F * T * C->DUT[cb] / den;
            b = bd2 * 2.0 * F * T * C->DUT[ca] / den;
          } else {
            a = bd2 * 2.0 * F * T * U * C->DUT[cb] / den;
            b = bd2 * 2.0 * F * T * U * C->DUT[ca] / den;
          }
        } else {
          const double den = C->V[ca] + C->V[cb] + EPS;
          if (with_ul) {
            a = bd2 * 2.0 * F * T * F * C->DUT[cb] / den;
            b = bd2 * 2.0 * F * T * F * C->DUT[ca] / den;
          } else {
            a = bd2 * 2.0 * F * T * U * F * C->DUT[cb] / den;
            b = bd2 * 2.0 * F * T * U * F * C->DUT[ca] / den;
          }
        }
        if (with_ul) { if (lower) u3[ca] = -a; else l3[ca] = -b; }
        L->d[cell] = L->d[cell] + (lower ? b : a);
      } else {
        const double FG = C->G3[ff], den = C->V[ca] + C->V[cb] + EPS;
        const double D = lower ? C->DUT[ca] : C->DUT[cb];
        if (with_ul) {
          if (lower) {
            u3[2 * ca] = -(bd2 * 2.0 * F * T * FG * C->DUT[cb] / den);
            u3[2 * ca + 1] = -(bd2 * 2.0 * F * FG * C->DUT[cb] / den);
          } else {
            l3[2 * ca] = -(bd2 * 2.0 * F * T * FG * C->DUT[ca] / den);
            l3[2 * ca + 1] = -(bd2 * 2.0 * F * FG * C->DUT[ca] / den);
          }
          L->d[4 * cell] = L->d[4 * cell] + bd2 * 2.0 * F * T * FG * D / den;
          L->d[4 * cell + 1] = L->d[4 * cell + 1] + bd2 * 2.0 * F * FG * D / den;
        } else {
          L->d[4 * cell] = L->d[4 * cell] + bd2 * 2.0 * F * T * U * FG * D / den;
          L->d[4 * cell + 1] = L->d[4 * cell + 1] + bd2 * 2.0 * F * U * FG * D / den;
        }
      }
    }
}

static void compute3d(const cctx *X, const cgrid *C, lev2 *L, double *u3, double *l3) {
  const int m3 = L->m3;
  plane2d(X, C, L, 0);
  for (int k = 0; k + 1 < m3; ++k) {
    yface(X, C, L, u3, l3, k, 1);
    plane2d(X, C, L, k + 1);
    yface(X, C, L, u3, l3, k, 0);
  }
}

/* ------------------------------------------------------------------- level set-up */
static void lev2_alloc(lev2 *L, int m1, int m2, int m3, int np) {
  const long n = (long)m1 * m2 * m3;
  L->m1 = m1; L->m2 = m2; L->m3 = m3;
  L->d = zalloc(n * np * np); L->u1 = zalloc(n * np * np); L->l1 = zalloc(n * np * np);
  L->u2 = zalloc(n * np); L->l2 = zalloc(n * np);
  const long p = (long)m1 * m2;
  L->res = zalloc(p * np);
  L->rc = zalloc((long)cdiv(m1, 2) * cdiv(m2, 2) * np);
  L->xc = zalloc((long)cdiv(m1, 2) * cdiv(m2, 2) * np);
  L->t = zalloc((long)m1 * np);
  L->w = zalloc((long)m1 * np * np);
}
static void lev2_free(lev2 *L) {
  free(L->d); free(L->u1); free(L->l1); free(L->u2); free(L->l2);
  free(L->res); free(L->rc); free(L->xc); free(L->t); free(L->w);
}

/* finest data in the reversed order from the reference-shaped inputs */
static void fine_grid(const orc_grid *g, const double *FU, const double *FV, const double *FW,
                      const double *V, const double *dx, const double *dy, const double *dz,
                      const double *DTU, const double *DTV, const double *DTW,
                      const double *DUT, cgrid *f) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const int n1 = nz, n2 = nx, n3 = ny;
  f->m1 = n1; f->m2 = n2; f->m3 = n3;
  const long nf1 = (long)(n1 + 1) * n2 * n3, nf2 = (long)n1 * (n2 + 1) * n3,
             nf3 = (long)n1 * n2 * (n3 + 1), nc = (long)n1 * n2 * n3;
  f->F1 = zalloc(nf1); f->T1 = zalloc(nf1); f->U1 = zalloc(nf1); f->G1 = zalloc(nf1);
  f->F2 = zalloc(nf2); f->T2 = zalloc(nf2); f->U2 = zalloc(nf2); f->G2 = zalloc(nf2);
  f->F3 = zalloc(nf3); f->T3 = zalloc(nf3); f->U3 = zalloc(nf3); f->G3 = zalloc(nf3);
  f->V = zalloc(nc); f->DUT = zalloc(nc);
  f->dd1 = zalloc(n1); f->dd2 = zalloc(n2); f->dd3 = zalloc(n3);
  for (int i = 0; i < n1; ++i) f->dd1[i] = dz[i];
  for (int j = 0; j < n2; ++j) f->dd2[j] = dx[j];
  for (int k = 0; k < n3; ++k) f->dd3[k] = dy[k];
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const long c = ix + (long)nx * (iy + (long)ny * iz);
        const long r = I3(iz, ix, iy, n1, n2);
        f->V[r] = V[c];
        f->DUT[r] = DUT[c];
      }
  /* FG of the interior faces (ReadWeights_Mod.f90:1376-1462); wall faces carry no
   * weight (closed walls), so their FG is never used */
  const int gf = g->grad_full;
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) { /* x faces -> F2 */
        const long s = ix + (long)(nx + 1) * (iy + (long)ny * iz), r = I2(iz, ix, iy, n1, n2);
        const double F = FU[s];
        f->F2[r] = F; f->T2[r] = DTU[s]; f->U2[r] = F / (F + EPS);
        if (ix > 0 && ix < nx) {
          const long a = (ix - 1) + (long)nx * (iy + (long)ny * iz);
          f->G2[r] = gf ? (V[a] + V[a + 1]) / (1.0 * (dx[ix] + dx[ix - 1] + EPS)) * F / (F + EPS) : F;
        }
      }
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) { /* y faces -> F3 */
        const long s = ix + (long)nx * (iy + (long)(ny + 1) * iz), r = I3(iz, ix, iy, n1, n2);
        const double F = FV[s];
        f->F3[r] = F; f->T3[r] = DTV[s]; f->U3[r] = F / (F + EPS);
        if (iy > 0 && iy < ny) {
          const long a = ix + (long)nx * ((iy - 1) + (long)ny * iz);
          f->G3[r] = gf ? (V[a] + V[a + nx]) / (1.0 * (dy[iy] + dy[iy - 1] + EPS)) * F / (F + EPS) : F;
        }
      }
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) { /* z faces -> F1 */
        const long s = ix + (long)nx * (iy + (long)ny * iz), r = I1(iz, ix, iy, n1, n2);
        const double F = FW[s];
        f->F1[r] = F; f->T1[r] = DTW[s]; f->U1[r] = F / (F + EPS);
        if (iz > 0 && iz < nz) {
          const long a = ix + (long)nx * (iy + (long)ny * (iz - 1));
          f->G1[r] = gf ? (V[a] + V[a + (long)nx * ny]) / (1.0 * (dz[iz] + dz[iz - 1] + EPS)) * F / (F + EPS) : F;
        }
      }
}

void orc_mg_default_control(orc_mg_control *c) {
  c->ref_level = 4;     /* INIT/Grid_Mod.f90:880, SRC/Block.grid */
  c->gauss_iter2 = 1;   /* MODULE/Control_Mod.f90:55-58 */
  c->gauss_iter3 = 1;
  c->mult_iter2 = 1;
  c->mult_iter3 = 1;
}

orc_mg *orc_mg_build(const orc_grid *g, const orc_mg_control *ctl, const double *FU,
                     const double *FV, const double *FW, const double *V, const double *dx,
                     const double *dy, const double *dz, const double *DTU, const double *DTV,
                     const double *DTW, const double *DUT) {
  if (!g || !ctl || ctl->ref_level < 1 || ctl->mult_iter2 < 1 || ctl->mult_iter3 < 1 ||
      ctl->gauss_iter2 < 0 || ctl->gauss_iter3 < 0)
    return NULL;
  orc_mg *M = (orc_mg *)calloc(1, sizeof(orc_mg));
  M->g = *g;
  M->ctl = *ctl;
  M->np = g->np1;
  M->n1 = g->nz; M->n2 = g->nx; M->n3 = g->ny;
  M->R = ctl->ref_level;
  const int tb = g->np1 == 2, np = M->np, R = M->R;
  cgrid f;
  memset(&f, 0, sizeof f);
  fine_grid(g, FU, FV, FW, V, dx, dy, dz, DTU, DTV, DTW, DUT, &f);
  cctx X = {g, np};
  M->L = (lev3 *)calloc(R, sizeof(lev3));
  for (int L3 = 0; L3 < R; ++L3) {
    lev3 *Q = M->L + L3;
    const int s3 = 1 << L3;
    cgrid C;
    memset(&C, 0, sizeof C);
    coarsen(&f, tb, s3, s3, s3, &C);
    Q->m1 = C.m1; Q->m2 = C.m2; Q->m3 = C.m3;
    Q->nq = R - L3;
    Q->q = (lev2 *)calloc(Q->nq, sizeof(lev2));
    const long n = (long)Q->m1 * Q->m2 * Q->m3;
    Q->u3 = zalloc(n * np); Q->l3 = zalloc(n * np);
    Q->res = zalloc(n * np);
    Q->rc = zalloc((long)cdiv(Q->m1, 2) * cdiv(Q->m2, 2) * cdiv(Q->m3, 2) * np);
    Q->xc = zalloc((long)cdiv(Q->m1, 2) * cdiv(Q->m2, 2) * cdiv(Q->m3, 2) * np);
    Q->t = zalloc((long)Q->m1 * Q->m2 * np);
    lev2_alloc(&Q->q[0], C.m1, C.m2, C.m3, np);
    compute3d(&X, &C, &Q->q[0], Q->u3, Q->l3);
    cgrid_free(&C);
    for (int L2 = L3 + 1; L2 < R; ++L2) {
      const int s12 = 1 << L2;
      cgrid D;
      memset(&D, 0, sizeof D);
      coarsen(&f, tb, s12, s12, s3, &D);
      lev2 *P = &Q->q[L2 - L3];
      lev2_alloc(P, D.m1, D.m2, D.m3, np);
      compute3d(&X, &D, P, NULL, NULL);
      cgrid_free(&D);
    }
  }
  cgrid_free(&f);
  const long nn = (long)M->n1 * M->n2 * M->n3 * np;
  M->xr = zalloc(nn);
  M->br = zalloc(nn);
  return M;
}

void orc_mg_free(orc_mg *M) {
  if (!M) return;
  for (int L3 = 0; L3 < M->R; ++L3) {
    lev3 *Q = M->L + L3;
    for (int q = 0; q < Q->nq; ++q) lev2_free(&Q->q[q]);
    free(Q->q); free(Q->u3); free(Q->l3); free(Q->res); free(Q->rc); free(Q->xc); free(Q->t);
  }
  free(M->L); free(M->xr); free(M->br);
  free(M);
}

/* ------------------------------------------------------------------ line and plane */
/* Solve_1D on line j of plane k of 2D level L, in place on b (length m1*np) */
static void solve1d(const lev2 *L, int np, int j, int k, double *b) {
  const int m1 = L->m1;
  const long base = (long)m1 * (j + (long)L->m2 * k);
  const double *d = L->d + base * np * np, *u = L->u1 + base * np * np,
               *l = L->l1 + base * np * np;
  double *w = L->w;
  if (np == 1) {
    double t = d[0];
    b[0] = b[0] / t;
    for (int i = 1; i < m1; ++i) {
      w[i - 1] = u[i - 1] / t;
      t = d[i] - l[i - 1] * w[i - 1];
      b[i] = (b[i] - l[i - 1] * b[i - 1]) / t;
    }
    for (int i = m1 - 2; i >= 0; --i) b[i] = b[i] - w[i] * b[i + 1];
  } else {
    double t[4], tmp[4], v[2];
    memcpy(t, d, sizeof t);
    div2(v, b, t); b[0] = v[0]; b[1] = v[1];
    for (int i = 1; i < m1; ++i) {
      div22(w + 4 * (i - 1), u + 4 * (i - 1), t);
      mul22(tmp, l + 4 * (i - 1), w + 4 * (i - 1));
      for (int q = 0; q < 4; ++q) t[q] = d[4 * i + q] - tmp[q];
      mul2(v, l + 4 * (i - 1), b + 2 * (i - 1));
      double rr[2] = {b[2 * i] - v[0], b[2 * i + 1] - v[1]};
      div2(v, rr, t);
      b[2 * i] = v[0]; b[2 * i + 1] = v[1];
    }
    for (int i = m1 - 2; i >= 0; --i) {
      mul2(v, w + 4 * i, b + 2 * (i + 1));
      b[2 * i] = b[2 * i] - v[0];
      b[2 * i + 1] = b[2 * i + 1] - v[1];
    }
  }
}

/* y = A x on plane k of 2D level L (MatVecT_1D/2D, MatVecTB_1D/2D); x, y planes */
static void matvec2(const lev2 *L, int np, int k, double *y, const double *x) {
  const int m1 = L->m1, m2 = L->m2;
  for (int j = 0; j < m2; ++j) {
    const long base = (long)m1 * (j + (long)m2 * k);
    const double *d = L->d + base * np * np, *u = L->u1 + base * np * np,
                 *l = L->l1 + base * np * np;
    const double *xj = x + (long)m1 * j * np;
    double *yj = y + (long)m1 * j * np;
    if (np == 1) {
      if (m1 > 1) {
        yj[0] = d[0] * xj[0] + u[0] * xj[1];
        for (int i = 1; i + 1 < m1; ++i)
          yj[i] = l[i - 1] * xj[i - 1] + d[i] * xj[i] + u[i] * xj[i + 1];
        yj[m1 - 1] = l[m1 - 2] * xj[m1 - 2] + d[m1 - 1] * xj[m1 - 1];
      } else {
        yj[0] = d[0] * xj[0];
      }
    } else {
      double a[2], b[2], c[2];
      if (m1 > 1) {
        mul2(a, d, xj); mul2(b, u, xj + 2);
        yj[0] = a[0] + b[0]; yj[1] = a[1] + b[1];
        for (int i = 1; i + 1 < m1; ++i) {
          mul2(a, l + 4 * (i - 1), xj + 2 * (i - 1));
          mul2(b, d + 4 * i, xj + 2 * i);
          mul2(c, u + 4 * i, xj + 2 * (i + 1));
          yj[2 * i] = a[0] + b[0] + c[0];
          yj[2 * i + 1] = a[1] + b[1] + c[1];
        }
        mul2(a, l + 4 * (m1 - 2), xj + 2 * (m1 - 2));
        mul2(b, d + 4 * (m1 - 1), xj + 2 * (m1 - 1));
        yj[2 * (m1 - 1)] = a[0] + b[0]; yj[2 * (m1 - 1) + 1] = a[1] + b[1];
      } else {
        mul2(a, d, xj);
        yj[0] = a[0]; yj[1] = a[1];
      }
    }
  }
  if (m2 > 1) {
    const long pb = (long)m1 * m2 * k;
    const double *u2 = L->u2 + pb * np, *l2 = L->l2 + pb * np;
    for (int j = 0; j < m2; ++j)
      for (int i = 0; i < m1; ++i) {
        const long c = i + (long)m1 * j;
        for (int r = 0; r < np; ++r) {
          double s = y[np * c + r];
          if (j == 0) {
            s = s + u2[np * c + r] * x[np * (c + m1)];
          } else if (j + 1 < m2) {
            s = s + l2[np * (c - m1) + r] * x[np * (c - m1)] + u2[np * c + r] * x[np * (c + m1)];
          } else {
            s = s + l2[np * (c - m1) + r] * x[np * (c - m1)];
          }
          y[np * c + r] = s;
        }
      }
  }
}

/* GaussSeidel_2D on plane k of level L: x (in/out) for b, gi symmetric sweeps */
static void gs2(const lev2 *L, int np, int k, const double *b, double *x, int gi) {
  const int m1 = L->m1, m2 = L->m2;
  const long pb = (long)m1 * m2 * k;
  const double *u2 = L->u2 + pb * np, *l2 = L->l2 + pb * np;
  double *t = L->t;
  const long ln = (long)m1 * np;
#define XL(j) (x + (long)(j) * ln)
#define BL(j) (b + (long)(j) * ln)
  if (m2 > 1) {
    for (int it = 0; it < gi; ++it) {
      memcpy(t, BL(0), sizeof(double) * ln);
      for (int j = 0; j + 1 < m2; ++j) {
        for (int i = 0; i < m1; ++i) {
          const long c = i + (long)m1 * j;
          for (int r = 0; r < np; ++r) XL(j)[np * i + r] = t[np * i + r] - u2[np * c + r] * XL(j + 1)[np * i];
        }
        solve1d(L, np, j, k, XL(j));
        for (int i = 0; i < m1; ++i) {
          const long c = i + (long)m1 * j;
          for (int r = 0; r < np; ++r) t[np * i + r] = BL(j + 1)[np * i + r] - l2[np * c + r] * XL(j)[np * i];
        }
      }
      memcpy(XL(m2 - 1), t, sizeof(double) * ln);
      solve1d(L, np, m2 - 1, k, XL(m2 - 1));
      for (int i = 0; i < m1; ++i) {
        const long c = i + (long)m1 * (m2 - 2);
        for (int r = 0; r < np; ++r) t[np * i + r] = BL(m2 - 2)[np * i + r] - u2[np * c + r] * XL(m2 - 1)[np * i];
      }
      for (int j = m2 - 2; j >= 1; --j) {
        for (int i = 0; i < m1; ++i) {
          const long c = i + (long)m1 * (j - 1);
          for (int r = 0; r < np; ++r) XL(j)[np * i + r] = t[np * i + r] - l2[np * c + r] * XL(j - 1)[np * i];
        }
        solve1d(L, np, j, k, XL(j));
        for (int i = 0; i < m1; ++i) {
          const long c = i + (long)m1 * (j - 1);
          for (int r = 0; r < np; ++r) t[np * i + r] = BL(j - 1)[np * i + r] - u2[np * c + r] * XL(j)[np * i];
        }
      }
    }
    memcpy(XL(0), t, sizeof(double) * ln);
    solve1d(L, np, 0, k, XL(0));
  } else {
    memcpy(XL(0), BL(0), sizeof(double) * ln);
    solve1d(L, np, 0, k, XL(0));
  }
#undef XL
#undef BL
}

/* RestrictT_2D (summation, DO i1; DO i2) / RestrictTB_2D (its explicit sums) */
static void restrict2(int tb, int np, int n1, int n2, const double *r, double *rc) {
  const int c1 = cdiv(n1, 2), c2 = cdiv(n2, 2);
#define R2(a, b) (r + np * ((long)(a) + (long)n1 * (b)))
#define C2(a, b) (rc + np * ((long)(a) + (long)c1 * (b)))
  if (!tb) {
    memset(rc, 0, sizeof(double) * np * c1 * c2);
    for (int i1 = 0; i1 < n1; ++i1)
      for (int i2 = 0; i2 < n2; ++i2)
        for (int q = 0; q < np; ++q) C2(i1 / 2, i2 / 2)[q] = C2(i1 / 2, i2 / 2)[q] + R2(i1, i2)[q];
  } else {
    for (int q = 0; q < np; ++q) C2(c1 - 1, c2 - 1)[q] = R2(n1 - 1, n2 - 1)[q];
    for (int iC = 0; iC < n1 / 2; ++iC)
      for (int q = 0; q < np; ++q) C2(iC, c2 - 1)[q] = R2(2 * iC, n2 - 1)[q] + R2(2 * iC + 1, n2 - 1)[q];
    for (int jC = 0; jC < n2 / 2; ++jC) {
      for (int q = 0; q < np; ++q) C2(c1 - 1, jC)[q] = R2(n1 - 1, 2 * jC)[q] + R2(n1 - 1, 2 * jC + 1)[q];
      for (int iC = 0; iC < n1 / 2; ++iC)
        for (int q = 0; q < np; ++q)
          C2(iC, jC)[q] = R2(2 * iC, 2 * jC)[q] + R2(2 * iC + 1, 2 * jC)[q] + R2(2 * iC, 2 * jC + 1)[q] +
                          R2(2 * iC + 1, 2 * jC + 1)[q];
    }
  }
#undef R2
#undef C2
}

/* ProlongateT_2D / ProlongateTB_2D: piecewise constant */
static void prolong2(int np, int n1, int n2, const double *xc, double *x) {
  const int c1 = cdiv(n1, 2);
  for (int i2 = 0; i2 < n2; ++i2)
    for (int i1 = 0; i1 < n1; ++i1)
      for (int q = 0; q < np; ++q)
        x[np * (i1 + (long)n1 * i2) + q] = xc[np * (i1 / 2 + (long)c1 * (i2 / 2)) + q];
}

/* Multigrid_2D of plane k over the 2D levels Q->q[0..nq-1] (b, x: planes of q[0]) */
static void mg2(const orc_mg *M, const lev3 *Q, int k, const double *b, double *x) {
  const int np = M->np, gi2 = M->ctl.gauss_iter2, nq = Q->nq, tb = np == 2;
  const double *bl[32];
  double *xl[32];
  bl[0] = b; xl[0] = x;
  int gi = 1, q = 0;
  for (;;) {
    const lev2 *L = &Q->q[q];
    gs2(L, np, k, bl[q], xl[q], gi);
    gi += gi2;
    if (q + 1 < nq) {
      const long n = (long)L->m1 * L->m2 * np;
      matvec2(L, np, k, L->res, xl[q]);
      for (long e = 0; e < n; ++e) L->res[e] = bl[q][e] - L->res[e];
      restrict2(tb, np, L->m1, L->m2, L->res, L->rc);
      bl[q + 1] = L->rc;
      xl[q + 1] = L->xc;
      memset(L->xc, 0, sizeof(double) * cdiv(L->m1, 2) * cdiv(L->m2, 2) * np);
      ++q;
    } else {
      gi -= gi2;
      break;
    }
  }
  for (;;) {
    const lev2 *L = &Q->q[q];
    gs2(L, np, k, bl[q], xl[q], gi);
    gi -= gi2;
    if (q > 0) {
      const lev2 *F = &Q->q[q - 1];
      prolong2(np, F->m1, F->m2, xl[q], F->res);
      const long n = (long)F->m1 * F->m2 * np;
      for (long e = 0; e < n; ++e) xl[q - 1][e] = xl[q - 1][e] + F->res[e];
      --q;
    } else {
      break;
    }
  }
}

/* Solve_2D: MultIter2 cycles */
static void solve2(const orc_mg *M, const lev3 *Q, int k, const double *b, double *x) {
  for (int it = 0; it < M->ctl.mult_iter2; ++it) mg2(M, Q, k, b, x);
}

/* ------------------------------------------------------------------------ 3D */
static void matvec3(const orc_mg *M, const lev3 *Q, double *y, const double *x) {
  const int np = M->np, m1 = Q->m1, m2 = Q->m2, m3 = Q->m3;
  const long pl = (long)m1 * m2 * np;
  for (int k = 0; k < m3; ++k) matvec2(&Q->q[0], np, k, y + k * pl, x + k * pl);
  if (m3 > 1) {
    const long pc = (long)m1 * m2;
    for (int k = 0; k < m3; ++k)
      for (long c = 0; c < pc; ++c) {
        const long e = c + pc * k;
        for (int r = 0; r < np; ++r) {
          double s = y[np * e + r];
          if (k == 0) s = s + Q->u3[np * e + r] * x[np * (e + pc)];
          else if (k + 1 < m3)
            s = s + Q->l3[np * (e - pc) + r] * x[np * (e - pc)] + Q->u3[np * e + r] * x[np * (e + pc)];
          else s = s + Q->l3[np * (e - pc) + r] * x[np * (e - pc)];
          y[np * e + r] = s;
        }
      }
  }
}

/* GaussSeidel_3D (TriDiagT_Mod.f90:951-1007, TriDiagTB_Mod.f90:818-874) */
static void gs3(const orc_mg *M, const lev3 *Q, const double *b, double *x, int gi) {
  const int np = M->np, m1 = Q->m1, m2 = Q->m2, m3 = Q->m3;
  const long pc = (long)m1 * m2, pl = pc * np;
  double *t = Q->t;
#define XP(k) (x + (long)(k) * pl)
#define BP(k) (b + (long)(k) * pl)
  if (m3 > 1) {
    for (int it = 0; it < gi; ++it) {
      memcpy(t, BP(0), sizeof(double) * pl);
      for (int k = 0; k + 1 < m3; ++k) {
        for (long c = 0; c < pc; ++c)
          for (int r = 0; r < np; ++r)
            t[np * c + r] = t[np * c + r] - Q->u3[np * (c + pc * k) + r] * XP(k + 1)[np * c];
        solve2(M, Q, k, t, XP(k));
        for (long c = 0; c < pc; ++c)
          for (int r = 0; r < np; ++r)
            t[np * c + r] = BP(k + 1)[np * c + r] - Q->l3[np * (c + pc * k) + r] * XP(k)[np * c];
      }
      solve2(M, Q, m3 - 1, t, XP(m3 - 1));
      for (long c = 0; c < pc; ++c)
        for (int r = 0; r < np; ++r)
          t[np * c + r] = BP(m3 - 2)[np * c + r] - Q->u3[np * (c + pc * (m3 - 2)) + r] * XP(m3 - 1)[np * c];
      for (int k = m3 - 2; k >= 1; --k) {
        for (long c = 0; c < pc; ++c)
          for (int r = 0; r < np; ++r)
            t[np * c + r] = t[np * c + r] - Q->l3[np * (c + pc * (k - 1)) + r] * XP(k - 1)[np * c];
        solve2(M, Q, k, t, XP(k));
        for (long c = 0; c < pc; ++c)
          for (int r = 0; r < np; ++r)
            t[np * c + r] = BP(k - 1)[np * c + r] - Q->u3[np * (c + pc * (k - 1)) + r] * XP(k)[np * c];
      }
    }
    solve2(M, Q, 0, t, XP(0));
    /* the extra solve of the top plane (TriDiagT_Mod.f90:998-1001) */
    for (long c = 0; c < pc; ++c)
      for (int r = 0; r < np; ++r)
        t[np * c + r] = BP(m3 - 1)[np * c + r] - Q->l3[np * (c + pc * (m3 - 2)) + r] * XP(m3 - 2)[np * c];
    solve2(M, Q, m3 - 1, t, XP(m3 - 1));
  } else {
    solve2(M, Q, 0, BP(0), XP(0));
  }
#undef XP
#undef BP
}

/* RestrictT_3D (DO i1; DO i2; DO i3) / RestrictTB_3D (its explicit sums) */
static void restrict3(int tb, int np, int n1, int n2, int n3, const double *r, double *rc) {
  const int c1 = cdiv(n1, 2), c2 = cdiv(n2, 2), c3 = cdiv(n3, 2);
#define R3(a, b, c) (r + np * ((long)(a) + (long)n1 * ((b) + (long)n2 * (c))))
#define C3(a, b, c) (rc + np * ((long)(a) + (long)c1 * ((b) + (long)c2 * (c))))
  if (!tb) {
    memset(rc, 0, sizeof(double) * np * c1 * c2 * c3);
    for (int i1 = 0; i1 < n1; ++i1)
      for (int i2 = 0; i2 < n2; ++i2)
        for (int i3 = 0; i3 < n3; ++i3)
          for (int q = 0; q < np; ++q)
            C3(i1 / 2, i2 / 2, i3 / 2)[q] = C3(i1 / 2, i2 / 2, i3 / 2)[q] + R3(i1, i2, i3)[q];
    return;
  }
  const int N1 = n1 - 1, N2 = n2 - 1, N3 = n3 - 1;
  for (int q = 0; q < np; ++q) C3(c1 - 1, c2 - 1, c3 - 1)[q] = R3(N1, N2, N3)[q];
  for (int iC = 0; iC < n1 / 2; ++iC)
    for (int q = 0; q < np; ++q)
      C3(iC, c2 - 1, c3 - 1)[q] = R3(2 * iC, N2, N3)[q] + R3(2 * iC + 1, N2, N3)[q];
  for (int jC = 0; jC < n2 / 2; ++jC) {
    const int jF = 2 * jC;
    for (int q = 0; q < np; ++q) C3(c1 - 1, jC, c3 - 1)[q] = R3(N1, jF, N3)[q] + R3(N1, jF + 1, N3)[q];
    for (int iC = 0; iC < n1 / 2; ++iC) {
      const int iF = 2 * iC;
      for (int q = 0; q < np; ++q)
        C3(iC, jC, c3 - 1)[q] = R3(iF, jF, N3)[q] + R3(iF + 1, jF, N3)[q] + R3(iF, jF + 1, N3)[q] +
                                R3(iF + 1, jF + 1, N3)[q];
    }
  }
  for (int kC = 0; kC < n3 / 2; ++kC) {
    const int kF = 2 * kC;
    for (int q = 0; q < np; ++q) C3(c1 - 1, c2 - 1, kC)[q] = R3(N1, N2, kF)[q] + R3(N1, N2, kF + 1)[q];
    for (int iC = 0; iC < n1 / 2; ++iC) {
      const int iF = 2 * iC;
      for (int q = 0; q < np; ++q)
        C3(iC, c2 - 1, kC)[q] = R3(iF, N2, kF)[q] + R3(iF + 1, N2, kF)[q] + R3(iF, N2, kF + 1)[q] +
                                R3(iF + 1, N2, kF + 1)[q];
    }
    for (int jC = 0; jC < n2 / 2; ++jC) {
      const int jF = 2 * jC;
      for (int q = 0; q < np; ++q)
        C3(c1 - 1, jC, kC)[q] = R3(N1, jF, kF)[q] + R3(N1, jF + 1, kF)[q] + R3(N1, jF, kF + 1)[q] +
                                R3(N1, jF + 1, kF + 1)[q];
      for (int iC = 0; iC < n1 / 2; ++iC) {
        const int iF = 2 * iC;
        for (int q = 0; q < np; ++q)
          C3(iC, jC, kC)[q] = R3(iF, jF, kF)[q] + R3(iF + 1, jF, kF)[q] + R3(iF, jF + 1, kF)[q] +
                              R3(iF, jF, kF + 1)[q] + R3(iF + 1, jF + 1, kF)[q] +
                              R3(iF + 1, jF, kF + 1)[q] + R3(iF, jF + 1, kF + 1)[q] +
                              R3(iF + 1, jF + 1, kF + 1)[q];
      }
    }
  }
#undef R3
#undef C3
}

static void prolong3(int np, int n1, int n2, int n3, const double *xc, double *x) {
  const int c1 = cdiv(n1, 2), c2 = cdiv(n2, 2);
  for (int i3 = 0; i3 < n3; ++i3)
    for (int i2 = 0; i2 < n2; ++i2)
      for (int i1 = 0; i1 < n1; ++i1)
        for (int q = 0; q < np; ++q)
          x[np * (i1 + (long)n1 * (i2 + (long)n2 * i3)) + q] =
              xc[np * (i1 / 2 + (long)c1 * (i2 / 2 + (long)c2 * (i3 / 2))) + q];
}

/* Multigrid_3D (TriDiagT_Mod.f90:1009-1071, TriDiagTB_Mod.f90:876-938) */
static void mg3(const orc_mg *M, const double *b, double *x) {
  const int np = M->np, gi3 = M->ctl.gauss_iter3, R = M->R, tb = np == 2;
  const double *bl[32];
  double *xl[32];
  bl[0] = b; xl[0] = x;
  int gi = 1, q = 0;
  for (;;) {
    const lev3 *Q = &M->L[q];
    gs3(M, Q, bl[q], xl[q], gi);
    gi += gi3;
    if (q + 1 < R) {
      const long n = (long)Q->m1 * Q->m2 * Q->m3 * np;
      matvec3(M, Q, Q->res, xl[q]);
      for (long e = 0; e < n; ++e) Q->res[e] = bl[q][e] - Q->res[e];
      restrict3(tb, np, Q->m1, Q->m2, Q->m3, Q->res, Q->rc);
      bl[q + 1] = Q->rc;
      xl[q + 1] = Q->xc;
      memset(Q->xc, 0, sizeof(double) * cdiv(Q->m1, 2) * cdiv(Q->m2, 2) * cdiv(Q->m3, 2) * np);
      ++q;
    } else {
      gi -= gi3;
      break;
    }
  }
  for (;;) {
    const lev3 *Q = &M->L[q];
    gs3(M, Q, bl[q], xl[q], gi);
    gi -= gi3;
    if (q > 0) {
      const lev3 *F = &M->L[q - 1];
      prolong3(np, F->m1, F->m2, F->m3, xl[q], F->res);
      const long n = (long)F->m1 * F->m2 * F->m3 * np;
      for (long e = 0; e < n; ++e) xl[q - 1][e] = xl[q - 1][e] + F->res[e];
      --q;
    } else {
      break;
    }
  }
}

/* ------------------------------------------------------------------ conversions, API */
/* reference order p(np, nx, ny, nz) <-> (np, z, x, y) */
static void to_rev(const orc_mg *M, const double *a, double *r) {
  const int nx = M->g.nx, ny = M->g.ny, nz = M->g.nz, np = M->np;
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const long c = ix + (long)nx * (iy + (long)ny * iz), e = iz + (long)nz * (ix + (long)nx * iy);
        for (int q = 0; q < np; ++q) r[np * e + q] = a[np * c + q];
      }
}
static void from_rev(const orc_mg *M, const double *r, double *a) {
  const int nx = M->g.nx, ny = M->g.ny, nz = M->g.nz, np = M->np;
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const long c = ix + (long)nx * (iy + (long)ny * iz), e = iz + (long)nz * (ix + (long)nx * iy);
        for (int q = 0; q < np; ++q) a[np * c + q] = r[np * e + q];
      }
}

void orc_mg_pressurepre(const orc_mg *M, double *x, const double *r) {
  orc_mg *W = (orc_mg *)M; /* work arrays only */
  const long n = (long)M->n1 * M->n2 * M->n3 * M->np;
  to_rev(M, r, W->br);
  memset(W->xr, 0, sizeof(double) * n);
  if (M->np == 1) {  /* VCycleCG, MultiTriT: Solve(AR, r, x) = MultIter3 cycles */
    for (int it = 0; it < M->ctl.mult_iter3; ++it) mg3(M, W->br, W->xr);
  } else {           /* MultiTriTB: one Multigrid_3D (MultIter = 1 there) */
    mg3(M, W->br, W->xr);
  }
  from_rev(M, W->xr, x);
}

int orc_mg_levels(const orc_mg *M, int level3, int *dims) {
  if (!M || level3 < 0 || level3 >= M->R) return -1;
  dims[0] = M->L[level3].m2; /* x */
  dims[1] = M->L[level3].m3; /* y */
  dims[2] = M->L[level3].m1; /* z */
  return M->L[level3].nq;
}

/* y = A x at 3D level `level3`, vectors in the reference order of that level's grid */
int orc_mg_apply_level(const orc_mg *M, int level3, double *y, const double *x) {
  if (!M || level3 < 0 || level3 >= M->R) return -1;
  const lev3 *Q = &M->L[level3];
  orc_mg tmp = *M;
  tmp.g.nx = Q->m2; tmp.g.ny = Q->m3; tmp.g.nz = Q->m1;
  const long n = (long)Q->m1 * Q->m2 * Q->m3 * M->np;
  double *xr = zalloc(n), *yr = zalloc(n);
  to_rev(&tmp, x, xr);
  matvec3(M, Q, yr, xr);
  from_rev(&tmp, yr, y);
  free(xr); free(yr);
  return 0;
}

/* restriction from 3D level `level3` to level3+1 (rc) and the prolongation back (x), in
 * the reference order of the two grids */
int orc_mg_restrict(const orc_mg *M, int level3, double *rc, const double *r) {
  if (!M || level3 < 0 || level3 + 1 >= M->R) return -1;
  const lev3 *Q = &M->L[level3], *C = &M->L[level3 + 1];
  orc_mg tf = *M, tc = *M;
  tf.g.nx = Q->m2; tf.g.ny = Q->m3; tf.g.nz = Q->m1;
  tc.g.nx = C->m2; tc.g.ny = C->m3; tc.g.nz = C->m1;
  const long nf = (long)Q->m1 * Q->m2 * Q->m3 * M->np, nc = (long)C->m1 * C->m2 * C->m3 * M->np;
  double *a = zalloc(nf), *b = zalloc(nc);
  to_rev(&tf, r, a);
  restrict3(M->np == 2, M->np, Q->m1, Q->m2, Q->m3, a, b);
  from_rev(&tc, b, rc);
  free(a); free(b);
  return 0;
}

int orc_mg_prolong(const orc_mg *M, int level3, double *x, const double *xc) {
  if (!M || level3 < 0 || level3 + 1 >= M->R) return -1;
  const lev3 *Q = &M->L[level3], *C = &M->L[level3 + 1];
  orc_mg tf = *M, tc = *M;
  tf.g.nx = Q->m2; tf.g.ny = Q->m3; tf.g.nz = Q->m1;
  tc.g.nx = C->m2; tc.g.ny = C->m3; tc.g.nz = C->m1;
  const long nf = (long)Q->m1 * Q->m2 * Q->m3 * M->np, nc = (long)C->m1 * C->m2 * C->m3 * M->np;
  double *a = zalloc(nc), *b = zalloc(nf);
  to_rev(&tc, xc, a);
  prolong3(M->np, Q->m1, Q->m2, Q->m3, a, b);
  from_rev(&tf, b, x);
  free(a); free(b);
  return 0;
}
