/*
 * asam_ensemble.h -- C-ABI of the batched (ensemble) sound stage solver on gfx950.
 *
 * An ensemble is a set of independent stage systems ("members") of uneven grid sizes,
 * each the system SolveSound solves for one block (reference SRC/SOLU/Iter_Mod.f90:9-30,
 * operator SRC/MULTI/JacAccGrav_Mod.f90:304-471), solved together on one GPU: every
 * kernel of the FP32-inner / FP64-refined solve covers all members in one launch (a
 * persistent grid walking a tile table that spans the members), the whole schedule of
 * one solve is captured once into a HIP graph and replayed with one graph launch, and
 * the tile shapes come from a built-in autotuner.
 *
 * The reference solves one block per call and has no batched entry point; each member
 * keeps SolveSound's conventions (operator MatVec1 = -A x, so the solve is (-A) x = b;
 * vectors p(1:np1, nx, ny, nz), component fastest; reference array shapes for the
 * geometry and state).  Ensemble vectors are the members' vectors concatenated in member
 * order; asam_ens_member_range gives each member's offset and length.
 *
 * Return codes are those of asam_sound.h.  A context is bound to the HIP device current
 * at creation and owns one stream; calls on one context are not re-entrant.
 */
#ifndef ASAM_ENSEMBLE_H
#define ASAM_ENSEMBLE_H

#include "asam_sound.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct asam_ens_ctx asam_ens_ctx;

/* Launch shapes of the batched kernels (chosen by asam_ens_autotune). */
typedef struct {
  int apply_bx;      /* stencil-apply tile width along x: 64, 128 or 256 (256 threads) */
  int apply_zsplit;  /* tiles per column along z (1..8): each member's nz split evenly */
  int apply_pf;      /* workgroups per CU of the persistent apply grid; 0: one per tile */
  int line_window;   /* levels per load window of the line solve: 4 or 8 */
  int line_lds;      /* LDS budget of the line solve's forward sweep per workgroup: 0 none
                        (output vector), 1 default (60 KiB), 2..160 that many KiB */
  int line_pf;       /* workgroups per CU of the persistent line grid; 0: one per tile */
  int blas_chunk;    /* vector elements per tile of the update kernels: 1024..65536 */
  int blas_pf;       /* workgroups per CU of the persistent update grid; 0: one per tile */
  int use_graph;     /* 1: capture the solve schedule into a HIP graph (default) */
  int apply_rows;    /* rows of a column each apply thread walks: 2 (default, the
                        neighbouring row from registers) or 1 (0 reads as 1) */
  int reserved[2];
} asam_ens_tuning;

/* Per-ensemble counters of the last solve. */
typedef struct {
  int graph_launches;   /* HIP graph launches */
  int graph_nodes;      /* nodes of the replayed graph (0 without one) */
  int eager_launches;   /* kernels enqueued outside the graph (schedule continuation) */
  int outer_steps;      /* refinement steps executed (maximum over members) */
  double ms;            /* device time of the whole solve (HIP events) */
} asam_ens_stats;

/* nmem members; grids[m] describes member m.  All members share np1 and grad_full. */
int asam_ens_create(asam_ens_ctx **ctx, int nmem, const asam_sound_grid *grids);
int asam_ens_destroy(asam_ens_ctx *ctx);
const char *asam_ens_last_error(const asam_ens_ctx *ctx);

/* Operator set-up of member m from geometry and state (the arguments and conventions
 * of asam_sound_setup). */
int asam_ens_setup(asam_ens_ctx *ctx, int m, const double *FU, const double *FV,
                   const double *FW, const double *VolC, const double *dx, const double *dy,
                   const double *dz, const double *rho, const double *rhotheta,
                   const double *T);

/* Total vector length (sum over members of np1*nx*ny*nz) and member m's range. */
long asam_ens_length(const asam_ens_ctx *ctx);
int asam_ens_member_range(const asam_ens_ctx *ctx, int m, long *offset, long *length);

/* Controls: method must be BiCGStab and precision ASAM_PREC_MIXED (inner_tol,
 * max_inner, max_outer as in asam_sound_control). */
int asam_ens_set_control(asam_ens_ctx *ctx, const asam_sound_control *ctl);

/* Solve every member to ||b - MatVec1(x)|| <= tol ||b - MatVec1(x0)|| (x0 = x on entry).
 * x, b: concatenated ensemble vectors on the host (asam_ens_solve) or on the device
 * (asam_ens_solve_device, device pointers of asam_ens_length doubles).  info: nmem
 * entries (may be NULL).  A member that does not converge within max_outer refinement
 * steps has status ASAM_NOT_CONVERGED; the call still returns 0. */
int asam_ens_solve(asam_ens_ctx *ctx, double *x, const double *b, double tol,
                   asam_sound_info *info);
int asam_ens_solve_device(asam_ens_ctx *ctx, double *x, const double *b, double tol,
                          asam_sound_info *info);

int asam_ens_get_tuning(const asam_ens_ctx *ctx, asam_ens_tuning *t);
int asam_ens_set_tuning(asam_ens_ctx *ctx, const asam_ens_tuning *t);
/* Launch-shape autotuner: for each candidate shape of the apply kernel, then of the line
 * solve, then of the update kernels (coordinate search), times whole FP32 inner BiCGStab
 * iterations over every member (all kernels and finalizes in solve order, reps iterations
 * each) and keeps the fastest.  ms[3] (may be NULL): ms per iteration of the chosen shapes
 * in ms[0], ms[1] = ms[2] = 0.  Requires every member set up; uses the solve's work
 * vectors. */
int asam_ens_autotune(asam_ens_ctx *ctx, int reps, asam_ens_tuning *best, double *ms);
int asam_ens_get_stats(const asam_ens_ctx *ctx, asam_ens_stats *stats);

/* Device memory on the context's GPU for callers without their own (ensemble vectors for
 * asam_ens_solve_device); copies run on the context's stream and complete before return. */
int asam_ens_dev_alloc(asam_ens_ctx *ctx, void **ptr, size_t bytes);
int asam_ens_dev_free(asam_ens_ctx *ctx, void *ptr);
int asam_ens_dev_upload(asam_ens_ctx *ctx, void *dst, const void *src, size_t bytes);
int asam_ens_dev_download(asam_ens_ctx *ctx, void *dst, const void *src, size_t bytes);
int asam_ens_dev_memset(asam_ens_ctx *ctx, void *dst, int value, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* ASAM_ENSEMBLE_H */
