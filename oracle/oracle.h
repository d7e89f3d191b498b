/*
 * oracle.h -- CPU restatement (FP64) of ASAM's sound stage solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library gives the reference answers the HIP product
 * path is tested against, and the CPU baseline timed beside it.  Nothing in the product
 * (asam_amd/, include/) links or calls it.
 *
 * Every function cites the reference routine it restates (paths relative to
 * the reference's SRC/ directory).
 */
#ifndef ASAM_ORACLE_H
#define ASAM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nx, ny, nz;   /* interior cells */
  int np1;          /* 1: MultiTriT (rho*theta), 2: MultiTriTB (rho*theta, rho) */
  int grad_full;    /* Control_Mod GradFull */
  double beta0_dt;  /* beta0*dtP */
  double grav;      /* GravComp (TB only) */
  double fac_anela; /* Physics_Mod FacAnela */
} orc_grid;

typedef struct orc_op orc_op;

/* Assemble the Neumann sound matrix from geometry and state
 * (JacAccGrav -> WeightDiagCompute 'Density' -> LaplTCompute / LaplTBCompute). */
orc_op *orc_assemble(const orc_grid *g, const double *FU, const double *FV, const double *FW,
                     const double *VolC, const double *dx, const double *dy, const double *dz,
                     const double *rho, const double *rhotheta, const double *T);
/* Same, but with the thermodynamic weights given directly (LaplTCompute inputs). */
orc_op *orc_assemble_weights(const orc_grid *g, const double *FU, const double *FV,
                             const double *FW, const double *VolC, const double *dx,
                             const double *dy, const double *dz, const double *DTU,
                             const double *DTV, const double *DTW, const double *DUT);
void orc_free(orc_op *A);
long orc_n(const orc_op *A);

/* Face weights computed by WeightDiagCompute ('Density'): theta at faces and DUT. */
void orc_weights(const orc_grid *g, const double *VolC, const double *rho,
                 const double *rhotheta, const double *T, double *DTU, double *DTV,
                 double *DTW, double *DUT);

/* Moist form (WeightDiagCompute 'Density', JacAccGrav_Mod.f90:4254-4270): vapour rhov,
 * cloud water rhol and rain rhor densities (any may be NULL); rho is the total density. */
void orc_weights_moist(const orc_grid *g, const double *VolC, const double *rho,
                       const double *rhotheta, const double *T, const double *rhov,
                       const double *rhol, const double *rhor, double *DTU, double *DTV,
                       double *DTW, double *DUT);

/* Export the nested tri-diagonal coefficients (x-fastest, face (c,c+e) stored at c).
 * T: d[N],ux,lx,uy,ly,uz,lz[N].  TB: d[4N] (2x2 column-major), uz,lz[4N], ux,lx,uy,ly[2N]. */
void orc_export(const orc_op *A, double *d, double *ux, double *lx, double *uy, double *ly,
                double *uz, double *lz);

void orc_apply(const orc_op *A, double *y, const double *x);     /* y = A x      (MatVecT_3D) */
void orc_matvec1(const orc_op *A, double *y, const double *x);   /* y = -A x     (MatVec1)    */
void orc_matvec1_T(const orc_op *A, double *y, const double *x); /* y = -A^T x              */
void orc_psolve1(const orc_op *A, double *y, const double *x);   /* y = -M^-1 x  (PSolve1)    */
void orc_psolve1_T(const orc_op *A, double *y, const double *x); /* y = -M^-T x             */
void orc_line_solve(const orc_op *A, double *y, const double *x); /* y = M^-1 x (Solve_1D)  */

/* Krylov solvers of the operator -A with preconditioner -M^-1.  transpose!=0 runs the
 * same sequence with -A^T and -M^-T (adjoint solve). */
/* BiCGStab (SpMatrix4CellPar_Mod.f90:245-348): maxiter/tol in-out with reference semantics. */
int orc_bicgstab(const orc_op *A, double *x, const double *b, int *maxiter, double *tol,
                 int transpose);
/* FGMRES(m) (Iter_Mod.f90:121-199): maxiter = restart cycles. Returns total Arnoldi steps. */
int orc_fgmres(const orc_op *A, double *x, const double *b, int maxiter, double tol, int m,
               int transpose, double *rel_res);
/* FGMRESV(m) (Iter_Mod.f90:201-288): FGMRES with classical Gram-Schmidt. */
int orc_fgmresv(const orc_op *A, double *x, const double *b, int maxiter, double tol, int m,
                int transpose, double *rel_res);
/* GMRES(m), left preconditioned (Iter_Mod.f90:32-119). Returns total Arnoldi steps. */
int orc_gmres(const orc_op *A, double *x, const double *b, int maxiter, double tol, int m,
              int transpose, double *rel_res);

double orc_dot(long n, const double *x, const double *y);
/* Right-hand side of the stage system (Rhs, MATRIX/Rhs_Mod.f90:110-157, one closed block,
 * dry 'Density' weights so DTT = 1): b1 = beta0dt * DivergenceS(F, theta_f, u)
 * (:46-82) - V * incr_rt (when fac_anela >= 1); TB: b2 = beta0dt * Divergence(F, u)
 * (:13-43) - V * incr_r.  U (nx+1,ny,nz), Vv (nx,ny+1,nz), W (nx,ny,nz+1) face
 * velocities; incr_rt, incr_r cell increments (incr_r unused for np1 = 1). */
void orc_rhs(const orc_grid *g, const double *FU, const double *FV, const double *FW,
             const double *V, const double *DTU, const double *DTV, const double *DTW,
             const double *U, const double *Vv, const double *W, const double *incr_rt,
             const double *incr_r, double *b);
/* Post-solve update (UpdateVelocity, MATRIX/Rhs_Mod.f90:589-952, one closed block, dry
 * 'Density': DUR = 0, DTT = 1, Cartesian metric): x1 *= DUT; TB: buoyancy correction of
 * the interior w faces with x2, then x2 *= DUR (= 0); pressure-gradient correction of the
 * interior faces (GradFull or Dual form); then incr_rt -= beta0dt DivergenceS(u)/V and
 * incr_r -= beta0dt Divergence(u)/V with the corrected velocities (incr_r may be NULL). */
void orc_update(const orc_grid *g, const double *FU, const double *FV, const double *FW,
                const double *V, const double *dx, const double *dy, const double *dz,
                const double *DTU, const double *DTV, const double *DTW, const double *DUT,
                double *x, double *U, double *Vv, double *W, double *incr_rt, double *incr_r);

void orc_set_threads(int n);
int orc_get_threads(void);

/* ---- the reference's own preconditioner, PressurePre (multigrid.c) -------------------
 * Semicoarsened multigrid with symmetric lexicographic y-plane Gauss-Seidel smoothing,
 * each plane solved by a 2D multigrid with symmetric x-line Gauss-Seidel smoothing and
 * exact z-line Thomas solves (MATRIX/PreCond_Mod.f90:599-622 -> VCycleCG,
 * MULTI/JacAccGrav_Mod.f90:4468-4517; hierarchy :685-1484; cycles MULTI/TriDiagT_Mod.f90,
 * TriDiagTB_Mod.f90). */
typedef struct {
  int ref_level;   /* RefLevel of the block (grid file; 4 in Block.grid), >= 1 */
  int gauss_iter2; /* Control_Mod GaussIter2 (default 1) */
  int gauss_iter3; /* Control_Mod GaussIter3 (default 1) */
  int mult_iter2;  /* Control_Mod MultIter2 (default 1) */
  int mult_iter3;  /* Control_Mod MultIter3 (default 1; MultiTriT only) */
} orc_mg_control;

typedef struct orc_mg orc_mg;

void orc_mg_default_control(orc_mg_control *c);
/* The level hierarchy from the geometry and the thermodynamic weights (the inputs of
 * orc_assemble_weights); NULL for an invalid control. */
orc_mg *orc_mg_build(const orc_grid *g, const orc_mg_control *ctl, const double *FU,
                     const double *FV, const double *FW, const double *VolC, const double *dx,
                     const double *dy, const double *dz, const double *DTU, const double *DTV,
                     const double *DTW, const double *DUT);
void orc_mg_free(orc_mg *M);
/* x = PressurePre(r): x = 0, then the V-cycles of VCycleCG; r, x in p(np1,nx,ny,nz) order. */
void orc_mg_pressurepre(const orc_mg *M, double *x, const double *r);
/* Grid of 3D level `level3` (dims = nx, ny, nz); returns the number of 2D levels of its
 * planes, -1 for a level that does not exist. */
int orc_mg_levels(const orc_mg *M, int level3, int *dims);
/* y = A x with the matrix of 3D level `level3` (MatVecT_3D / MatVecTB_3D). */
int orc_mg_apply_level(const orc_mg *M, int level3, double *y, const double *x);
/* Residual restriction from level3 to level3+1 (RestrictT_3D / RestrictTB_3D) and the
 * prolongation back (ProlongateT_3D / ProlongateTB_3D). */
int orc_mg_restrict(const orc_mg *M, int level3, double *rc, const double *r);
int orc_mg_prolong(const orc_mg *M, int level3, double *x, const double *xc);
/* Use M (NULL: the vertical-line solve) as the preconditioner of A's PSolve1 and hence of
 * the Krylov solvers above (forward solves only; the transposed solves keep the line
 * preconditioner). */
void orc_set_pressurepre(orc_op *A, const orc_mg *M);

#ifdef __cplusplus
}
#endif
#endif
