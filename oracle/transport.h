/*
 * transport.h -- CPU restatement (FP64) of ASAM's transport stage solve: BiCGStab over
 * the SpDiag transport Jacobian with the Gauss-Seidel preconditioner.
 *
 * TEST INFRASTRUCTURE ONLY.  The reference answers the HIP transport path is tested
 * against, and the CPU baseline timed beside it; nothing in the product (asam_amd/,
 * include/) links or calls it.
 *
 * Layout (as in the reference, paths relative to its SRC/ directory):
 *   - one block with a one-cell halo: m = n = (nx+2)(ny+2)(nz+2) entries, x fastest
 *     (MODULE/Vector4Cell_Mod.f90 SpDiagAllocate :1020-1063);
 *   - a SpDiag matrix is Val(m, NumDiag) column-major with offsets DiagPtr(1:NumDiag) and
 *     main diagonal Diag; entry Val(i,d) is A(i - DiagPtr(d), i) (CG/MatSpDiag_Mod.f90
 *     SpAVec_SpDiag :95-115).  The transport matrices JacTMom/JacTPot have the 7 offsets
 *     -(nx+2)(ny+2), -(nx+2), -1, 0, 1, nx+2, (nx+2)(ny+2) and Diag = 4; JacFall has the
 *     offsets 0, (nx+2)(ny+2) and Diag = 1;
 *   - a vector of ncomp components is ncomp consecutive haloed arrays (Vector4Cell_T
 *     Vec(ic)%c(ix0:ix1+1, iy0:iy1+1, iz0:iz1+1, 1));
 *   - the per-cell LU factors (JacSChemLU / JacSMetLU after GefaMatrix) are ncomp*ncomp
 *     arrays over the interior cells, entry (ic,kc) at lu[(ic*ncomp + kc)*ncell + q].
 */
#ifndef ASAM_ORACLE_TRANSPORT_H
#define ASAM_ORACLE_TRANSPORT_H

#ifdef __cplusplus
extern "C" {
#endif

/* Operator of one component (PSolveSp / MatVecMetSp branches, MODULE/Vector4Cell_Mod.f90
 * :605-644, 694-740). */
#define ORC_TR_MOM 0  /* JacTMom */
#define ORC_TR_POT 1  /* JacTPot */
#define ORC_TR_FALL 2 /* JacTPot + alpha*JacFall (ScaleMat2(ic) = alpha < 0) */

typedef struct {
  int n, numdiag, diag; /* diag: 1-based index of the main diagonal */
  const int *ptr;
  const double *val;
} orc_spdiag;

/* y = A x (+ alpha B x when B != NULL): SpAVec_SpDiag / SpABVec_SpDiag :95-143, with B's
 * values scaled by alpha first, as MatVecMetSp does (Vector4Cell_Mod.f90:626-629). */
void orc_tr_spavec(const orc_spdiag *A, const orc_spdiag *B, double alpha, double *y,
                   const double *x);
/* x = GS(b): GaussSeidel_SpDiag :190-235 (C == NULL) / GaussSeidel_SpDiagAB :237-296 with
 * C scaled by alpha; iterg = IterG (reference 5).  x and b may alias. */
void orc_tr_gauss_seidel(const orc_spdiag *A, const orc_spdiag *C, double alpha, int iterg,
                         double *x, const double *b);
/* PSolveMatrix (Vector4Cell_Mod.f90:742-776) on the interior of the ncomp haloed arrays v. */
void orc_tr_lu_solve(int nx, int ny, int nz, int ncomp, const double *lu, double *v);
/* GefaMatrix (Vector4Cell_Mod.f90:1404-1429): lu = LU(I - fac*J) in place, no pivoting. */
void orc_tr_gefa(long ncell, int ncomp, double fac, double *lu);

typedef struct orc_tr orc_tr;

/* tmom, tpot: Val of the 7-diagonal matrices (m x 7, column-major); fall: Val of the
 * 2-diagonal JacFall (m x 2) or NULL; lu: per-cell LU factors or NULL (identity);
 * op_mv / op_ps: per-component operator (ORC_TR_*) of MatVec and PSolve; alpha: per
 * component ScaleMat2 (used with ORC_TR_FALL).  The arrays are referenced, not copied. */
orc_tr *orc_tr_create(int nx, int ny, int nz, int ncomp, int iterg, const double *tmom,
                      const double *tpot, const double *fall, const double *lu,
                      const int *op_mv, const int *op_ps, const double *alpha);
void orc_tr_free(orc_tr *T);
long orc_tr_m(const orc_tr *T);
/* MatVecChem_Vec4CellSp / MatVecMet_Vec4CellSp (MODULE/SpMatrix4CellPar_Mod.f90:56-84):
 * y = x + LU^-1 (T x - x) on the interior, T x on the halo.  init != 0 takes JacTPot for
 * every component (the MatVecChem form BiCGStab uses for its first residual, :164,271). */
void orc_tr_matvec(const orc_tr *T, double *y, const double *x, int init);
/* PSolve_Vec4CellSp (:86-102) -> PSolveSp (Vector4Cell_Mod.f90:694-740): per-component
 * Gauss-Seidel, then the six halo faces set to zero. */
void orc_tr_psolve(const orc_tr *T, double *y, const double *x);
/* Dot_VecCell -> Dot_VectorCell (Vector4Cell_Mod.f90:1120-1132): interior cells only. */
double orc_tr_dot(const orc_tr *T, const double *x, const double *y);
/* BiCGStabChem_Vec4CellSp / BiCGStabMet_Vec4CellSp (SpMatrix4CellPar_Mod.f90:140-348).
 * maxiter, tol: in/out with the reference's conventions (-Iter on the half-step exit,
 * Iter on the full-step exit, 2*Iter when not converged; tol returns ||r||; rho == 0
 * returns maxiter = Iter, tol = 0).  Returns 0, or 2 on the rho == 0 exit. */
int orc_tr_bicgstab(const orc_tr *T, double *x, const double *b, int *maxiter, double *tol);

#ifdef __cplusplus
}
#endif
#endif
