/*
 * oracle.c -- CPU restatement (FP64) of ASAM's implicit sound stage solve.
 *
 * TEST INFRASTRUCTURE ONLY: the reference answers the HIP library is tested against,
 * and the CPU baseline timed beside it.  The product path never links or calls this file.
 *
 * What is restated (reference paths relative to SRC/):
 *   - face weights:  MULTI/JacAccGrav_Mod.f90 IniJacAccGrav :286-302 (DUU),
 *                    WeightDiagCompute 'Density' :4254-4274 (DUT),
 *                    CellToFaceRhoWeight :3314-3321 (volume-weighted face theta)
 *   - gradient factor FG: INIT/ReadWeights_Mod.f90:1376-1462
 *   - Neumann matrix: MULTI/TriDiagT_Mod.f90 ComputeDual_1D/2D/3D :32-60,253-317,693-787,
 *                     ComputeFull_1D/2D/3D :62-90,319-399,789-900, Allocate_1D :92-111;
 *                     MULTI/TriDiagTB_Mod.f90 Compute_1D/2D/3D :32-93,261-343,663-780,
 *                     Allocate_1D :95-119.  Boundary factors are zero (SetFactorN,
 *                     MULTI/Reverse_Mod.f90:215-297), so wall faces add nothing.
 *   - apply:         MatVecT_1D/2D/3D (TriDiagT_Mod.f90:113-133,401-428,902-936),
 *                    MatVecTB_1D/2D/3D (TriDiagTB_Mod.f90:121-141,345-372,782-816);
 *                    MatVec1 = -A x (MATRIX/MatVec_Mod.f90:220-232)
 *   - line solve:    Solve_1D (TriDiagT_Mod.f90:135-155; TriDiagTB_Mod.f90:143-163 with
 *                    CG/Matrix_Mod.f90 Mult_22/Mult_2/Div_22/Div_2 :34-94)
 *   - Krylov:        BiCGStab SpMatrix4CellPar_Mod.f90:245-348;
 *                    FGMRES / FGMRESV / GMRES SOLU/Iter_Mod.f90:121-199 / :201-288 /
 *                    :32-119
 *
 * The preconditioner is the vertical-line block-Jacobi: every z-column's
 * tri-diagonal (or 2x2 block tri-diagonal) line matrix of the assembled
 * operator solved exactly by Solve_1D; PSolve1 semantics (returns -M^-1 x,
 * MATRIX/PreCond_Mod.f90:1555-1571 with b1pBTx2 negation).
 *
 * Cell index c = i + nx*(j + ny*k) (x fastest, Fortran order).  Face (c, c+e)
 * along direction e is stored at c.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EPS 1.0e-30
#define RD (1004.0 - 717.0)
#define CPD 1004.0
#define RV 461.0   /* Parameter_Mod.f90:25-29 */
#define CPV 1885.0
#define CPL 4186.0

struct orc_op {
  orc_grid g;
  long N;
  double *d, *ux, *lx, *uy, *ly, *uz, *lz;
  const orc_mg *mg; /* PressurePre of the reference (multigrid.c) when set */
};

void orc_set_pressurepre(orc_op *A, const orc_mg *M) { A->mg = M; }

long orc_n(const orc_op *A) { return A->N * A->g.np1; }

void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ weights */
void orc_weights(const orc_grid *g, const double *V, const double *rho, const double *rth,
                 const double *T, double *DTU, double *DTV, double *DTW, double *DUT) {
  orc_weights_moist(g, V, rho, rth, T, NULL, NULL, NULL, DTU, DTV, DTW, DUT);
}

void orc_weights_moist(const orc_grid *g, const double *V, const double *rho,
                       const double *rth, const double *T, const double *rhov,
                       const double *rhol, const double *rhor, double *DTU, double *DTV,
                       double *DTW, double *DUT) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const long N = (long)nx * ny * nz;
  double *th = (double *)malloc(sizeof(double) * N);
  for (long c = 0; c < N; ++c) th[c] = rth[c] / (rho[c] + EPS);
  /* DUT, JacAccGrav_Mod.f90:4258-4268: RhoLLoc = RhoL + RhoR, RhoD = Rho - RhoV - RhoL + Eps,
   * Rm = Rd + Rv RhoV/RhoD, Cpml = Cpd + Cpv RhoV/RhoD + Cpl RhoL/RhoD, Kappa = Rm/Cpml */
  for (long c = 0; c < N; ++c) {
    const double rv = rhov ? rhov[c] : 0.0;
    const double rl = (rhol ? rhol[c] : 0.0) + (rhor ? rhor[c] : 0.0);
    const double rhoD = rho[c] - rv - rl + EPS;
    const double Rm = RD + RV * rv / rhoD;
    const double Cpml = CPD + CPV * rv / rhoD + CPL * rl / rhoD;
    const double kappa = Rm / Cpml;
    DUT[c] = T[c] / (rth[c] + EPS) * (rhoD * RD + rv * RV) / (1.0 - kappa);
  }
  /* Face theta: volume-weighted mean of the two cells (JacAccGrav_Mod.f90:3316-3319);
   * wall faces take the adjacent cell's theta (they carry zero weight, SetFactorN). */
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        long f = i + (long)(nx + 1) * (j + (long)ny * k);
        long cl = (i - 1) + (long)nx * (j + (long)ny * k), cr = cl + 1;
        if (i == 0) DTU[f] = th[cr];
        else if (i == nx) DTU[f] = th[cl];
        else DTU[f] = (th[cr] * V[cr] + th[cl] * V[cl]) / (V[cr] + V[cl] + EPS);
      }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i < nx; ++i) {
        long f = i + (long)nx * (j + (long)(ny + 1) * k);
        long cl = i + (long)nx * ((j - 1) + (long)ny * k), cr = cl + nx;
        if (j == 0) DTV[f] = th[cr];
        else if (j == ny) DTV[f] = th[cl];
        else DTV[f] = (th[cr] * V[cr] + th[cl] * V[cl]) / (V[cr] + V[cl] + EPS);
      }
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        long f = i + (long)nx * (j + (long)ny * k);
        long cl = i + (long)nx * (j + (long)ny * (k - 1)), cr = cl + (long)nx * ny;
        if (k == 0) DTW[f] = th[cr];
        else if (k == nz) DTW[f] = th[cl];
        else DTW[f] = (th[cr] * V[cr] + th[cl] * V[cl]) / (V[cr] + V[cl] + EPS);
      }
  free(th);
}

/* --------------------------------------------------------------- assembly */
static orc_op *alloc_op(const orc_grid *g) {
  orc_op *A = (orc_op *)calloc(1, sizeof(orc_op));
  A->g = *g;
  A->N = (long)g->nx * g->ny * g->nz;
  const long N = A->N;
  const int w = g->np1 == 1 ? 1 : 4, wh = g->np1 == 1 ? 1 : 2;
  A->d = (double *)calloc((size_t)N * w, sizeof(double));
  A->uz = (double *)calloc((size_t)N * w, sizeof(double));
  A->lz = (double *)calloc((size_t)N * w, sizeof(double));
  A->ux = (double *)calloc((size_t)N * wh, sizeof(double));
  A->lx = (double *)calloc((size_t)N * wh, sizeof(double));
  A->uy = (double *)calloc((size_t)N * wh, sizeof(double));
  A->ly = (double *)calloc((size_t)N * wh, sizeof(double));
  /* Allocate_1D: d = 1e-40 (TriDiagT_Mod.f90:101; TriDiagTB_Mod.f90:104-109) */
  for (long c = 0; c < N; ++c) {
    if (g->np1 == 1) {
      A->d[c] = 1.0e-40;
    } else {
      A->d[4 * c + 0] = 1.0e-40;
      A->d[4 * c + 1] = 0.0;
      A->d[4 * c + 2] = 0.0;
      A->d[4 * c + 3] = 1.0e-40;
    }
  }
  return A;
}

/* Interior face coupling of the MultiTriT operator between cells a and b=a+e.
 * coef  = (beta0*dtP)^2 * 2 * F*DTU*DUU * [F/(Va+Vb+Eps)  (dual)  |  1/(da+db) (full)]
 * u(a) = -coef*DUT(b), l(a) = -coef*DUT(a), d(a) += coef*DUT(a), d(b) += coef*DUT(b)
 * (ComputeDual_1D :48-53, ComputeFull_1D :78-83; 2D/3D identical per direction). */
static void face_T(orc_op *A, long a, long b, double *u, double *l, double F, double DTU,
                   double da, double db, const double *Vol, const double *DUT) {
  const orc_grid *g = &A->g;
  const double bd = g->beta0_dt;
  const double DUU = F / (F + EPS);
  double ca, cb;
  if (g->grad_full) {
    /* Metr1*Metr2 = 1 on the Cartesian grid (ReadWeights_Mod.f90:1193) */
    ca = (bd * bd) * 2.0 * F * DTU * DUU * DUT[a] / ((da + db) * 1.0 * 1.0);
    cb = (bd * bd) * 2.0 * F * DTU * DUU * DUT[b] / ((da + db) * 1.0 * 1.0);
  } else {
    ca = (bd * bd) * 2.0 * F * DTU * DUU * F * DUT[a] / (Vol[a] + Vol[b] + EPS);
    cb = (bd * bd) * 2.0 * F * DTU * DUU * F * DUT[b] / (Vol[a] + Vol[b] + EPS);
  }
  u[a] = -cb;
  l[a] = -ca;
  A->d[a] += ca;
  A->d[b] += cb;
}

/* Interior face coupling of the MultiTriTB operator (horizontal directions):
 * 2-vectors acting on component 1 of the neighbour (TriDiagTB_Mod.f90:287-303). */
static void face_TB_h(orc_op *A, long a, long b, double *u, double *l, double F, double FG,
                      double DTU, const double *Vol, const double *DUT) {
  const orc_grid *g = &A->g;
  const double bd = g->beta0_dt;
  const double DUU = F / (F + EPS);
  const double den = Vol[a] + Vol[b] + EPS;
  double c1a = (bd * bd) * 2.0 * F * DTU * DUU * FG * DUT[a] / den;
  double c1b = (bd * bd) * 2.0 * F * DTU * DUU * FG * DUT[b] / den;
  double c2a = (bd * bd) * 2.0 * F * DUU * FG * DUT[a] / den;
  double c2b = (bd * bd) * 2.0 * F * DUU * FG * DUT[b] / den;
  u[2 * a + 0] = -c1b;
  u[2 * a + 1] = -c2b;
  l[2 * a + 0] = -c1a;
  l[2 * a + 1] = -c2a;
  A->d[4 * a + 0] += c1a; /* d(1,1) */
  A->d[4 * a + 1] += c2a; /* d(2,1) */
  A->d[4 * b + 0] += c1b;
  A->d[4 * b + 1] += c2b;
}

/* Vertical face of the MultiTriTB operator: full 2x2 blocks with gravity
 * (TriDiagTB_Mod.f90 Compute_1D :51-82). Column-major 2x2: [11,21,12,22]. */
static void face_TB_z(orc_op *A, long a, long b, double F, double FG, double DTU,
                      const double *Vol, const double *DUT) {
  const orc_grid *g = &A->g;
  const double bd = g->beta0_dt, G = g->grav;
  const double DUU = F / (F + EPS);
  const double den = Vol[a] + Vol[b] + EPS;
  double *u = A->uz + 4 * a, *l = A->lz + 4 * a, *da = A->d + 4 * a, *db = A->d + 4 * b;
  u[0] = -(bd * bd) * 2.0 * F * DTU * DUU * FG * DUT[b] / den;
  u[1] = -(bd * bd) * 2.0 * F * DUU * FG * DUT[b] / den;
  l[0] = -(bd * bd) * 2.0 * F * DTU * DUU * FG * DUT[a] / den;
  l[1] = -(bd * bd) * 2.0 * F * DUU * FG * DUT[a] / den;
  u[2] = u[2] - G * (bd * bd) * F * DTU * Vol[b] / den;
  u[3] = u[3] - G * (bd * bd) * F * Vol[b] / den;
  l[2] = l[2] + G * (bd * bd) * F * DTU * Vol[a] / den;
  l[3] = l[3] + G * (bd * bd) * F * Vol[a] / den;
  da[0] += (bd * bd) * 2.0 * F * DTU * DUU * FG * DUT[a] / den;
  da[1] += (bd * bd) * 2.0 * F * DUU * FG * DUT[a] / den;
  db[0] += (bd * bd) * 2.0 * F * DTU * DUU * FG * DUT[b] / den;
  db[1] += (bd * bd) * 2.0 * F * DUU * FG * DUT[b] / den;
  da[2] -= G * (bd * bd) * F * DTU * Vol[a] / den;
  da[3] -= G * (bd * bd) * F * Vol[a] / den;
  db[2] += G * (bd * bd) * F * DTU * Vol[b] / den;
  db[3] += G * (bd * bd) * F * Vol[b] / den;
}

/* FG (ReadWeights_Mod.f90:1376-1462): GradFull -> (Va+Vb)/(Metr*(da+db+Eps))*F/(F+Eps), else F */
static double fg_of(const orc_grid *g, double F, double Va, double Vb, double da, double db) {
  if (g->grad_full) return (Va + Vb) / (1.0 * (db + da + EPS)) * F / (F + EPS);
  return F;
}

orc_op *orc_assemble_weights(const orc_grid *g, const double *FU, const double *FV,
                             const double *FW, const double *V, const double *dx,
                             const double *dy, const double *dz, const double *DTU,
                             const double *DTV, const double *DTW, const double *DUT) {
  orc_op *A = alloc_op(g);
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const long sy = nx, sz = (long)nx * ny;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        long c = i + sy * j + sz * k;
        if (i + 1 < nx) { /* x-face i+1 */
          long f = (i + 1) + (long)(nx + 1) * (j + (long)ny * k);
          if (g->np1 == 1)
            face_T(A, c, c + 1, A->ux, A->lx, FU[f], DTU[f], dx[i], dx[i + 1], V, DUT);
          else
            face_TB_h(A, c, c + 1, A->ux, A->lx, FU[f],
                      fg_of(g, FU[f], V[c], V[c + 1], dx[i], dx[i + 1]), DTU[f], V, DUT);
        }
        if (j + 1 < ny) { /* y-face j+1 */
          long f = i + (long)nx * ((j + 1) + (long)(ny + 1) * k);
          if (g->np1 == 1)
            face_T(A, c, c + sy, A->uy, A->ly, FV[f], DTV[f], dy[j], dy[j + 1], V, DUT);
          else
            face_TB_h(A, c, c + sy, A->uy, A->ly, FV[f],
                      fg_of(g, FV[f], V[c], V[c + sy], dy[j], dy[j + 1]), DTV[f], V, DUT);
        }
        if (k + 1 < nz) { /* z-face k+1 */
          long f = i + (long)nx * (j + (long)ny * (k + 1));
          if (g->np1 == 1)
            face_T(A, c, c + sz, A->uz, A->lz, FW[f], DTW[f], dz[k], dz[k + 1], V, DUT);
          else
            face_TB_z(A, c, c + sz, FW[f], fg_of(g, FW[f], V[c], V[c + sz], dz[k], dz[k + 1]),
                      DTW[f], V, DUT);
        }
      }
  /* d += FacAnela*Vol (ComputeDual_1D :56-58); TB: d(2,2) += Vol (TriDiagTB :88-91) */
  for (long c = 0; c < A->N; ++c) {
    if (g->np1 == 1) {
      A->d[c] += g->fac_anela * V[c];
    } else {
      A->d[4 * c + 0] += g->fac_anela * V[c];
      A->d[4 * c + 3] += V[c];
    }
  }
  return A;
}

orc_op *orc_assemble(const orc_grid *g, const double *FU, const double *FV, const double *FW,
                     const double *V, const double *dx, const double *dy, const double *dz,
                     const double *rho, const double *rth, const double *T) {
  const long nx = g->nx, ny = g->ny, nz = g->nz;
  double *DTU = (double *)malloc(sizeof(double) * (nx + 1) * ny * nz);
  double *DTV = (double *)malloc(sizeof(double) * nx * (ny + 1) * nz);
  double *DTW = (double *)malloc(sizeof(double) * nx * ny * (nz + 1));
  double *DUT = (double *)malloc(sizeof(double) * nx * ny * nz);
  orc_weights(g, V, rho, rth, T, DTU, DTV, DTW, DUT);
  orc_op *A = orc_assemble_weights(g, FU, FV, FW, V, dx, dy, dz, DTU, DTV, DTW, DUT);
  free(DTU); free(DTV); free(DTW); free(DUT);
  return A;
}

void orc_free(orc_op *A) {
  if (!A) return;
  free(A->d); free(A->ux); free(A->lx); free(A->uy); free(A->ly); free(A->uz); free(A->lz);
  free(A);
}

void orc_export(const orc_op *A, double *d, double *ux, double *lx, double *uy, double *ly,
                double *uz, double *lz) {
  const long N = A->N;
  const size_t w = A->g.np1 == 1 ? 1 : 4, wh = A->g.np1 == 1 ? 1 : 2;
  memcpy(d, A->d, sizeof(double) * N * w);
  memcpy(uz, A->uz, sizeof(double) * N * w);
  memcpy(lz, A->lz, sizeof(double) * N * w);
  memcpy(ux, A->ux, sizeof(double) * N * wh);
  memcpy(lx, A->lx, sizeof(double) * N * wh);
  memcpy(uy, A->uy, sizeof(double) * N * wh);
  memcpy(ly, A->ly, sizeof(double) * N * wh);
}

/* ----------------------------------------------------------------- apply */
/* y = A x  (MatVecT_3D / MatVecTB_3D); transpose!=0 gives y = A^T x. */
static void apply_impl(const orc_op *A, double *y, const double *x, int transpose) {
  const orc_grid *g = &A->g;
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const long sy = nx, sz = (long)nx * ny;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        long c = i + sy * j + sz * k;
        if (g->np1 == 1) {
          double s = A->d[c] * x[c];
          if (!transpose) {
            if (i > 0) s += A->lx[c - 1] * x[c - 1];
            if (i + 1 < nx) s += A->ux[c] * x[c + 1];
            if (j > 0) s += A->ly[c - sy] * x[c - sy];
            if (j + 1 < ny) s += A->uy[c] * x[c + sy];
            if (k > 0) s += A->lz[c - sz] * x[c - sz];
            if (k + 1 < nz) s += A->uz[c] * x[c + sz];
          } else {
            if (i > 0) s += A->ux[c - 1] * x[c - 1];
            if (i + 1 < nx) s += A->lx[c] * x[c + 1];
            if (j > 0) s += A->uy[c - sy] * x[c - sy];
            if (j + 1 < ny) s += A->ly[c] * x[c + sy];
            if (k > 0) s += A->uz[c - sz] * x[c - sz];
            if (k + 1 < nz) s += A->lz[c] * x[c + sz];
          }
          y[c] = s;
        } else {
          const double *d = A->d + 4 * c;
          const double *xc = x + 2 * c;
          double y1, y2;
          if (!transpose) {
            /* Mult_2(d, x): y1 = d11 x1 + d12 x2, y2 = d21 x1 + d22 x2 */
            y1 = d[0] * xc[0] + d[2] * xc[1];
            y2 = d[1] * xc[0] + d[3] * xc[1];
            if (k > 0) {
              const double *l = A->lz + 4 * (c - sz), *xm = x + 2 * (c - sz);
              y1 += l[0] * xm[0] + l[2] * xm[1];
              y2 += l[1] * xm[0] + l[3] * xm[1];
            }
            if (k + 1 < nz) {
              const double *u = A->uz + 4 * c, *xp = x + 2 * (c + sz);
              y1 += u[0] * xp[0] + u[2] * xp[1];
              y2 += u[1] * xp[0] + u[3] * xp[1];
            }
            if (i > 0) { y1 += A->lx[2 * (c - 1)] * x[2 * (c - 1)]; y2 += A->lx[2 * (c - 1) + 1] * x[2 * (c - 1)]; }
            if (i + 1 < nx) { y1 += A->ux[2 * c] * x[2 * (c + 1)]; y2 += A->ux[2 * c + 1] * x[2 * (c + 1)]; }
            if (j > 0) { y1 += A->ly[2 * (c - sy)] * x[2 * (c - sy)]; y2 += A->ly[2 * (c - sy) + 1] * x[2 * (c - sy)]; }
            if (j + 1 < ny) { y1 += A->uy[2 * c] * x[2 * (c + sy)]; y2 += A->uy[2 * c + 1] * x[2 * (c + sy)]; }
          } else {
            /* d^T x */
            y1 = d[0] * xc[0] + d[1] * xc[1];
            y2 = d[2] * xc[0] + d[3] * xc[1];
            if (k > 0) { /* (A^T)_{c,c-z} = (A_{c-z,c})^T = u(c-z)^T */
              const double *u = A->uz + 4 * (c - sz), *xm = x + 2 * (c - sz);
              y1 += u[0] * xm[0] + u[1] * xm[1];
              y2 += u[2] * xm[0] + u[3] * xm[1];
            }
            if (k + 1 < nz) { /* (A^T)_{c,c+z} = l(c)^T */
              const double *l = A->lz + 4 * c, *xp = x + 2 * (c + sz);
              y1 += l[0] * xp[0] + l[1] * xp[1];
              y2 += l[2] * xp[0] + l[3] * xp[1];
            }
            /* horizontal: A_{(c,r),(c+1,1)} = ux(c)[r]; A_{(c+1,r),(c,1)} = lx(c)[r] */
            if (i > 0) y1 += A->ux[2 * (c - 1)] * x[2 * (c - 1)] + A->ux[2 * (c - 1) + 1] * x[2 * (c - 1) + 1];
            if (i + 1 < nx) y1 += A->lx[2 * c] * x[2 * (c + 1)] + A->lx[2 * c + 1] * x[2 * (c + 1) + 1];
            if (j > 0) y1 += A->uy[2 * (c - sy)] * x[2 * (c - sy)] + A->uy[2 * (c - sy) + 1] * x[2 * (c - sy) + 1];
            if (j + 1 < ny) y1 += A->ly[2 * c] * x[2 * (c + sy)] + A->ly[2 * c + 1] * x[2 * (c + sy) + 1];
          }
          y[2 * c] = y1;
          y[2 * c + 1] = y2;
        }
      }
}

void orc_apply(const orc_op *A, double *y, const double *x) { apply_impl(A, y, x, 0); }

void orc_matvec1(const orc_op *A, double *y, const double *x) {
  apply_impl(A, y, x, 0);
  const long n = orc_n(A);
  for (long c = 0; c < n; ++c) y[c] = -y[c];
}

void orc_matvec1_T(const orc_op *A, double *y, const double *x) {
  apply_impl(A, y, x, 1);
  const long n = orc_n(A);
  for (long c = 0; c < n; ++c) y[c] = -y[c];
}

/* ------------------------------------------------------------ line solve */
/* Solve_1D (TriDiagT_Mod.f90:135-155) on one column of length n, stride s. */
static void thomas_T(int n, long s, const double *d, const double *u, const double *l,
                     double *b, double *w) {
  double t = d[0];
  b[0] = b[0] / t;
  for (int i = 1; i < n; ++i) {
    w[i - 1] = u[(i - 1) * s] / t;
    t = d[i * s] - l[(i - 1) * s] * w[i - 1];
    b[i * s] = (b[i * s] - l[(i - 1) * s] * b[(i - 1) * s]) / t;
  }
  for (int i = n - 2; i >= 0; --i) b[i * s] = b[i * s] - w[i] * b[(i + 1) * s];
}

/* CG/Matrix_Mod.f90:34-94 (2x2, column-major [11,21,12,22]) */
static void mult22(double *r, const double *b, const double *c) {
  double t0 = b[0] * c[0] + b[2] * c[1];
  double t2 = b[0] * c[2] + b[2] * c[3];
  double t1 = b[1] * c[0] + b[3] * c[1];
  double t3 = b[1] * c[2] + b[3] * c[3];
  r[0] = t0; r[1] = t1; r[2] = t2; r[3] = t3;
}
static void mult2(double *r, const double *b, const double *c) {
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

/* Solve_1D of MultiTriTB (TriDiagTB_Mod.f90:143-163). Blocks at 4*i*s, vectors at 2*i*s. */
static void thomas_TB(int n, long s, const double *d, const double *u, const double *l,
                      double *b, double *w) {
  double t[4], tmp[4], v[2];
  memcpy(t, d, sizeof t);
  div2(v, b, t); b[0] = v[0]; b[1] = v[1];
  for (int i = 1; i < n; ++i) {
    div22(w + 4 * (i - 1), u + 4 * (i - 1) * s, t);
    mult22(tmp, l + 4 * (i - 1) * s, w + 4 * (i - 1));
    for (int q = 0; q < 4; ++q) t[q] = d[4 * i * s + q] - tmp[q];
    mult2(v, l + 4 * (i - 1) * s, b + 2 * (i - 1) * s);
    double rr[2] = {b[2 * i * s] - v[0], b[2 * i * s + 1] - v[1]};
    div2(v, rr, t);
    b[2 * i * s] = v[0]; b[2 * i * s + 1] = v[1];
  }
  for (int i = n - 2; i >= 0; --i) {
    mult2(v, w + 4 * i, b + 2 * (i + 1) * s);
    b[2 * i * s] -= v[0];
    b[2 * i * s + 1] -= v[1];
  }
}

static void transpose22(double *r, const double *a) { r[0] = a[0]; r[1] = a[2]; r[2] = a[1]; r[3] = a[3]; }

static void line_solve_impl(const orc_op *A, double *y, const double *x, int transpose) {
  const orc_grid *g = &A->g;
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const long sz = (long)nx * ny;
  const long n = orc_n(A);
  memcpy(y, x, sizeof(double) * n);
  if (g->np1 == 1) {
#pragma omp parallel
    {
      double *w = (double *)malloc(sizeof(double) * (nz + 1));
#pragma omp for schedule(static)
      for (long col = 0; col < sz; ++col) {
        /* transposed line: sub-diagonal u, super-diagonal l */
        if (!transpose) thomas_T(nz, sz, A->d + col, A->uz + col, A->lz + col, y + col, w);
        else thomas_T(nz, sz, A->d + col, A->lz + col, A->uz + col, y + col, w);
      }
      free(w);
    }
  } else {
#pragma omp parallel
    {
      double *w = (double *)malloc(sizeof(double) * 4 * (nz + 1));
      double *dt = NULL, *ut = NULL, *lt = NULL;
      if (transpose) {
        dt = (double *)malloc(sizeof(double) * 4 * nz);
        ut = (double *)malloc(sizeof(double) * 4 * nz);
        lt = (double *)malloc(sizeof(double) * 4 * nz);
      }
#pragma omp for schedule(static)
      for (long col = 0; col < sz; ++col) {
        if (!transpose) {
          thomas_TB(nz, sz, A->d + 4 * col, A->uz + 4 * col, A->lz + 4 * col, y + 2 * col, w);
        } else {
          for (int k = 0; k < nz; ++k) {
            transpose22(dt + 4 * k, A->d + 4 * (col + k * sz));
            transpose22(ut + 4 * k, A->lz + 4 * (col + k * sz));
            transpose22(lt + 4 * k, A->uz + 4 * (col + k * sz));
          }
          double *yc = (double *)malloc(sizeof(double) * 2 * nz);
          for (int k = 0; k < nz; ++k) { yc[2 * k] = y[2 * (col + k * sz)]; yc[2 * k + 1] = y[2 * (col + k * sz) + 1]; }
          thomas_TB(nz, 1, dt, ut, lt, yc, w);
          for (int k = 0; k < nz; ++k) { y[2 * (col + k * sz)] = yc[2 * k]; y[2 * (col + k * sz) + 1] = yc[2 * k + 1]; }
          free(yc);
        }
      }
      free(w); free(dt); free(ut); free(lt);
    }
  }
}

void orc_line_solve(const orc_op *A, double *y, const double *x) { line_solve_impl(A, y, x, 0); }

void orc_psolve1(const orc_op *A, double *y, const double *x) {
  const long n = orc_n(A);
  double *t = (double *)malloc(sizeof(double) * n);
  for (long c = 0; c < n; ++c) t[c] = -x[c]; /* b1pBTx2: b%p := -x%p */
  if (A->mg) orc_mg_pressurepre(A->mg, y, t); /* PressurePre(b, b) */
  else line_solve_impl(A, y, t, 0);
  free(t);
}

void orc_psolve1_T(const orc_op *A, double *y, const double *x) {
  const long n = orc_n(A);
  double *t = (double *)malloc(sizeof(double) * n);
  for (long c = 0; c < n; ++c) t[c] = -x[c];
  line_solve_impl(A, y, t, 1);
  free(t);
}

/* ---------------------------------------------------------------- BLAS-1 */
double orc_dot(long n, const double *x, const double *y) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}
static void axpy(long n, double a, const double *x, double *y) { /* y = y + a x */
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) y[i] = y[i] + a * x[i];
}
static void xpay(long n, const double *x, double a, double *y) { /* y = x + a y */
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) y[i] = x[i] + a * y[i];
}
static void scal(long n, double a, double *y) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) y[i] = a * y[i];
}

static void op(const orc_op *A, double *y, const double *x, int tr) {
  if (tr) orc_matvec1_T(A, y, x); else orc_matvec1(A, y, x);
}
static void pre(const orc_op *A, double *y, const double *x, int tr) {
  if (tr) orc_psolve1_T(A, y, x); else orc_psolve1(A, y, x);
}

/* --------------------------------------------------------------- BiCGStab */
int orc_bicgstab(const orc_op *A, double *x, const double *b, int *maxiter, double *tol,
                 int tr) {
  const long n = orc_n(A);
  double *r = (double *)malloc(sizeof(double) * n), *rt = (double *)malloc(sizeof(double) * n);
  double *rh = (double *)malloc(sizeof(double) * n), *p = (double *)malloc(sizeof(double) * n);
  double *ph = (double *)malloc(sizeof(double) * n), *t = (double *)malloc(sizeof(double) * n);
  double *v = (double *)malloc(sizeof(double) * n);
  double alpha = 0.0, beta, omega = 0.0, rho = 0.0, rhoOld, rNorm, rNorm0;
  const int MaxIter = *maxiter;
  const double Tol = *tol;
  int status = 1;
  op(A, r, x, tr);                  /* r = Op x */
  xpay(n, b, -1.0, r);              /* r = b - r */
  memcpy(rt, r, sizeof(double) * n);
  rNorm0 = sqrt(orc_dot(n, r, r));
  for (int iter = 1; iter <= MaxIter; ++iter) {
    rhoOld = rho;
    rho = orc_dot(n, rt, r);
    if (rho == 0.0) { *maxiter = iter; *tol = 0.0; status = -1; break; }
    if (iter == 1) {
      memcpy(p, r, sizeof(double) * n);
    } else {
      beta = (rho / rhoOld) * (alpha / omega);
      axpy(n, -omega, v, p);
      xpay(n, r, beta, p);
    }
    pre(A, ph, p, tr);
    op(A, v, ph, tr);
    alpha = rho / orc_dot(n, rt, v);
    axpy(n, -alpha, v, r);
    rNorm = sqrt(orc_dot(n, r, r));
    if (rNorm <= Tol * rNorm0) {
      axpy(n, alpha, ph, x);
      *maxiter = -iter; *tol = rNorm; status = 0; break;
    }
    pre(A, rh, r, tr);
    op(A, t, rh, tr);
    omega = orc_dot(n, t, r) / orc_dot(n, t, t);
    axpy(n, alpha, ph, x);
    axpy(n, omega, rh, x);
    axpy(n, -omega, t, r);
    rNorm = sqrt(orc_dot(n, r, r));
    if (rNorm <= Tol * rNorm0) { *maxiter = iter; *tol = rNorm; status = 0; break; }
    if (iter == MaxIter) { *maxiter = 2 * iter; *tol = rNorm; status = 1; break; }
  }
  free(r); free(rt); free(rh); free(p); free(ph); free(t); free(v);
  return status;
}

/* ----------------------------------------------------------------- FGMRES */
#define MAXM 64
/* FGMRES (Iter_Mod.f90:121-199) orthogonalises each new Arnoldi vector by modified
 * Gram-Schmidt: H(k,i) is taken after the projections on v_1..v_k-1 are removed.
 * FGMRESV (Iter_Mod.f90:201-288) is the same iteration with classical Gram-Schmidt: every
 * H(k,i) is the product of v_k with the same copy w of the new vector (:238-241), then all
 * projections are removed in order k = 1..i (:242-244).  For one closed block its
 * Exchange(w) calls are copies and its 'Rho' print is output only. */
static int fgmres_run(const orc_op *A, double *x, const double *b, int maxiter, double tol,
                      int m, int tr, double *rel_res, int classical) {
  const long n = orc_n(A);
  if (m > MAXM) m = MAXM;
  double **v = (double **)malloc(sizeof(double *) * (m + 1));
  double **z = (double **)malloc(sizeof(double *) * m);
  for (int i = 0; i <= m; ++i) v[i] = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < m; ++i) z[i] = (double *)malloc(sizeof(double) * n);
  double H[MAXM + 1][MAXM], c[MAXM], s[MAXM], bHat[MAXM + 1], y[MAXM];
  double beta, delta, rho = 0.0, rho0 = 0.0, temp;
  int nR = 0, total = 0;
  for (int iter = 1; iter <= maxiter; ++iter) {
    op(A, v[0], x, tr);
    xpay(n, b, -1.0, v[0]);        /* DXPBY(b,-1,v1): v1 = b - v1 */
    beta = sqrt(orc_dot(n, v[0], v[0]));
    if (iter == 1) rho0 = beta;
    if (beta <= 0.0) { rho = 0.0; break; }
    scal(n, 1.0 / beta, v[0]);
    bHat[0] = beta * 1.0;
    for (int i = 0; i < m; ++i) {
      pre(A, z[i], v[i], tr);
      op(A, v[i + 1], z[i], tr);
      ++total;
      if (classical) {
        for (int k = 0; k <= i; ++k) H[k][i] = orc_dot(n, v[k], v[i + 1]);
        for (int k = 0; k <= i; ++k) axpy(n, -H[k][i], v[k], v[i + 1]);
      } else {
        for (int k = 0; k <= i; ++k) {
          H[k][i] = orc_dot(n, v[k], v[i + 1]);
          axpy(n, -H[k][i], v[k], v[i + 1]);
        }
      }
      H[i + 1][i] = sqrt(orc_dot(n, v[i + 1], v[i + 1]));
      scal(n, 1.0 / H[i + 1][i], v[i + 1]);
      for (int k = 1; k <= i; ++k) {
        temp = c[k - 1] * H[k - 1][i] + s[k - 1] * H[k][i];
        H[k][i] = -s[k - 1] * H[k - 1][i] + c[k - 1] * H[k][i];
        H[k - 1][i] = temp;
      }
      delta = sqrt(H[i][i] * H[i][i] + H[i + 1][i] * H[i + 1][i]);
      c[i] = H[i][i] / delta;
      s[i] = H[i + 1][i] / delta;
      H[i][i] = c[i] * H[i][i] + s[i] * H[i + 1][i];
      bHat[i + 1] = -s[i] * bHat[i];
      bHat[i] = c[i] * bHat[i];
      rho = fabs(bHat[i + 1]);
      if (rho <= tol * rho0) { nR = i + 1; break; }
      nR = m;
    }
    y[nR - 1] = bHat[nR - 1] / H[nR - 1][nR - 1];
    for (int k = nR - 2; k >= 0; --k) {
      y[k] = bHat[k];
      for (int i = k + 1; i < nR; ++i) y[k] = y[k] - H[k][i] * y[i];
      y[k] = y[k] / H[k][k];
    }
    for (int i = 0; i < nR; ++i) axpy(n, y[i], z[i], x);
    if (rho <= tol * rho0) break;
  }
  if (rel_res) *rel_res = rho0 > 0.0 ? rho / rho0 : 0.0;
  for (int i = 0; i <= m; ++i) free(v[i]);
  for (int i = 0; i < m; ++i) free(z[i]);
  free(v); free(z);
  return total;
}

int orc_fgmres(const orc_op *A, double *x, const double *b, int maxiter, double tol, int m,
               int tr, double *rel_res) {
  return fgmres_run(A, x, b, maxiter, tol, m, tr, rel_res, 0);
}

int orc_fgmresv(const orc_op *A, double *x, const double *b, int maxiter, double tol, int m,
                int tr, double *rel_res) {
  return fgmres_run(A, x, b, maxiter, tol, m, tr, rel_res, 1);
}

/* ------------------------------------------------------------------ GMRES */
int orc_gmres(const orc_op *A, double *x, const double *b, int maxiter, double tol, int m,
              int tr, double *rel_res) {
  const long n = orc_n(A);
  if (m > MAXM) m = MAXM;
  double **v = (double **)malloc(sizeof(double *) * (m + 1));
  for (int i = 0; i <= m; ++i) v[i] = (double *)malloc(sizeof(double) * n);
  double *w = (double *)malloc(sizeof(double) * n);
  double H[MAXM + 1][MAXM], c[MAXM], s[MAXM], bHat[MAXM + 1], y[MAXM];
  double beta, delta, rho = 0.0, rho0 = 0.0, temp;
  int nR = 0, total = 0;
  for (int iter = 1; iter <= maxiter; ++iter) {
    op(A, w, x, tr);
    xpay(n, b, -1.0, w);
    pre(A, v[0], w, tr);
    beta = sqrt(orc_dot(n, v[0], v[0]));
    if (iter == 1) rho0 = beta;
    if (beta <= 0.0) { rho = 0.0; break; }
    scal(n, 1.0 / beta, v[0]);
    bHat[0] = beta * 1.0;
    for (int i = 0; i < m; ++i) {
      op(A, w, v[i], tr);
      pre(A, v[i + 1], w, tr);
      ++total;
      for (int k = 0; k <= i; ++k) {
        H[k][i] = orc_dot(n, v[k], v[i + 1]);
        axpy(n, -H[k][i], v[k], v[i + 1]);
      }
      H[i + 1][i] = sqrt(orc_dot(n, v[i + 1], v[i + 1]));
      scal(n, 1.0 / H[i + 1][i], v[i + 1]);
      for (int k = 1; k <= i; ++k) {
        temp = c[k - 1] * H[k - 1][i] + s[k - 1] * H[k][i];
        H[k][i] = -s[k - 1] * H[k - 1][i] + c[k - 1] * H[k][i];
        H[k - 1][i] = temp;
      }
      delta = sqrt(H[i][i] * H[i][i] + H[i + 1][i] * H[i + 1][i]);
      c[i] = H[i][i] / delta;
      s[i] = H[i + 1][i] / delta;
      H[i][i] = c[i] * H[i][i] + s[i] * H[i + 1][i];
      bHat[i + 1] = -s[i] * bHat[i];
      bHat[i] = c[i] * bHat[i];
      rho = fabs(bHat[i + 1]);
      if (rho <= tol * rho0) { nR = i + 1; break; }
      nR = m;
    }
    y[nR - 1] = bHat[nR - 1] / H[nR - 1][nR - 1];
    for (int k = nR - 2; k >= 0; --k) {
      y[k] = bHat[k];
      for (int i = k + 1; i < nR; ++i) y[k] = y[k] - H[k][i] * y[i];
      y[k] = y[k] / H[k][k];
    }
    for (int i = 0; i < nR; ++i) axpy(n, y[i], v[i], x);
    if (rho <= tol * rho0) break;
  }
  if (rel_res) *rel_res = rho0 > 0.0 ? rho / rho0 : 0.0;
  for (int i = 0; i <= m; ++i) free(v[i]);
  free(v); free(w);
  return total;
}


/* ------------------------------------------------------------------ caller side */
/* Face-flux divergence of one cell (DivergenceS / Divergence, Rhs_Mod.f90:13-82): the
 * reference's term order, west, east, south, north, bottom, top; th* NULL: no weight. */
static double div_cell(const orc_grid *g, int i, int j, int k, const double *FU,
                       const double *FV, const double *FW, const double *thu,
                       const double *thv, const double *thw, const double *U,
                       const double *Vv, const double *W) {
  const long nx = g->nx, ny = g->ny;
  const long uw = i + (nx + 1) * (j + ny * k), ue = uw + 1;
  const long vs = i + nx * (j + (ny + 1) * k), vn = vs + nx;
  const long wb = i + nx * (j + ny * k), wt = wb + nx * ny;
  double s;
  if (thu) {
    s = -FU[uw] * thu[uw] * U[uw];
    s += FU[ue] * thu[ue] * U[ue];
    s += -FV[vs] * thv[vs] * Vv[vs];
    s += FV[vn] * thv[vn] * Vv[vn];
    s += -FW[wb] * thw[wb] * W[wb];
    s += FW[wt] * thw[wt] * W[wt];
  } else {
    s = -FU[uw] * U[uw];
    s += FU[ue] * U[ue];
    s += -FV[vs] * Vv[vs];
    s += FV[vn] * Vv[vn];
    s += -FW[wb] * W[wb];
    s += FW[wt] * W[wt];
  }
  return s;
}

void orc_rhs(const orc_grid *g, const double *FU, const double *FV, const double *FW,
             const double *V, const double *DTU, const double *DTV, const double *DTW,
             const double *U, const double *Vv, const double *W, const double *incr_rt,
             const double *incr_r, double *b) {
  const int nx = g->nx, ny = g->ny, nz = g->nz, np1 = g->np1;
  const double bd = g->beta0_dt;
#pragma omp parallel for collapse(2)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long c = i + (long)nx * (j + (long)ny * k);
        double b1 = bd * div_cell(g, i, j, k, FU, FV, FW, DTU, DTV, DTW, U, Vv, W);
        if (g->fac_anela >= 1.0) b1 -= V[c] * incr_rt[c];
        b[np1 * c] = b1;
        if (np1 == 2)
          b[2 * c + 1] = bd * div_cell(g, i, j, k, FU, FV, FW, NULL, NULL, NULL, U, Vv, W) -
                         V[c] * incr_r[c];
      }
}

void orc_update(const orc_grid *g, const double *FU, const double *FV, const double *FW,
                const double *V, const double *dx, const double *dy, const double *dz,
                const double *DTU, const double *DTV, const double *DTW, const double *DUT,
                double *x, double *U, double *Vv, double *W, double *incr_rt, double *incr_r) {
  const int nx = g->nx, ny = g->ny, nz = g->nz, np1 = g->np1;
  const long N = (long)nx * ny * nz, sy = nx, sz = (long)nx * ny;
  const double bd = g->beta0_dt;
  /* x1 = x1*DUT (:611) */
  for (long c = 0; c < N; ++c) x[np1 * c] = x[np1 * c] * DUT[c];
  if (np1 == 2) {
    /* buoyancy correction of the interior w faces (:613-621), then x2 = x2*DUR = 0 */
    for (int k = 0; k + 1 < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const long c = i + sy * j + sz * k, u = c + sz;
          const double t = bd * g->grav * (V[c] * x[2 * c + 1] + V[u] * x[2 * u + 1]) /
                           (V[u] + V[c] + EPS);
          W[c + sz] = W[c + sz] - t;
        }
    for (long c = 0; c < N; ++c) x[2 * c + 1] = x[2 * c + 1] * 0.0;
  }
  /* interior faces: u = u - grad p (GradFull :679-745 / Dual :771-838, Cartesian metric) */
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long c = i + sy * j + sz * k;
        const double p1 = x[np1 * c], p2 = np1 == 2 ? x[2 * c + 1] : 0.0;
        if (i + 1 < nx) {
          const long f = (i + 1) + (long)(nx + 1) * (j + (long)ny * k), n = c + 1;
          const double duu = FU[f] / (FU[f] + EPS);
          double dp = x[np1 * n] - p1;
          if (np1 == 2) dp = dp + (x[2 * n + 1] - p2);
          const double t = g->grad_full ? 2.0 * bd * duu * dp / (dx[i + 1] + dx[i])
                                        : 2.0 * bd * FU[f] * duu * dp / (V[n] + V[c] + EPS);
          U[f] = U[f] - t;
        }
        if (j + 1 < ny) {
          const long f = i + (long)nx * ((j + 1) + (long)(ny + 1) * k), n = c + sy;
          const double duv = FV[f] / (FV[f] + EPS);
          double dp = x[np1 * n] - p1;
          if (np1 == 2) dp = dp + (x[2 * n + 1] - p2);
          const double t = g->grad_full ? 2.0 * bd * duv * dp / (dy[j + 1] + dy[j])
                                        : 2.0 * bd * FV[f] * duv * dp / (V[n] + V[c] + EPS);
          Vv[f] = Vv[f] - t;
        }
        if (k + 1 < nz) {
          const long f = c + sz, n = c + sz;
          const double duw = FW[f] / (FW[f] + EPS);
          double dp = x[np1 * n] - p1;
          if (np1 == 2) dp = dp + (x[2 * n + 1] - p2);
          if (g->grad_full) {
            const double t = 2.0 * bd * duw * dp / (dz[k + 1] + dz[k]);
            /* MultiTriT scales the w correction by FW/(FW+Eps) (:705-713) */
            W[f] = np1 == 1 ? W[f] - FW[f] / (FW[f] + EPS) * t : W[f] - t;
          } else {
            W[f] = W[f] - 2.0 * bd * FW[f] * duw * dp / (V[n] + V[c] + EPS);
          }
        }
      }
  /* increments from the corrected face fluxes (:840-873) */
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long c = i + sy * j + sz * k;
        const long uw = i + (long)(nx + 1) * (j + (long)ny * k), ue = uw + 1;
        const long vs = i + (long)nx * (j + (long)(ny + 1) * k), vn = vs + nx;
        const long wb = c, wt = c + sz;
        double s = FU[ue] * DTU[ue] * U[ue];
        s += -FU[uw] * DTU[uw] * U[uw];
        s += FV[vn] * DTV[vn] * Vv[vn];
        s += -FV[vs] * DTV[vs] * Vv[vs];
        s += FW[wt] * DTW[wt] * W[wt];
        s += -FW[wb] * DTW[wb] * W[wb];
        incr_rt[c] = incr_rt[c] - bd * s / (V[c] + EPS);
        if (incr_r) {
          double r = FU[ue] * U[ue];
          r += -FU[uw] * U[uw];
          r += FV[vn] * Vv[vn];
          r += -FV[vs] * Vv[vs];
          r += FW[wt] * W[wt];
          r += -FW[wb] * W[wb];
          incr_r[c] = incr_r[c] - bd * r / (V[c] + EPS);
        }
      }
}
