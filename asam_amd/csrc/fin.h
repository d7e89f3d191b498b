// fin.h -- scalar tails of the Krylov iterations: the single-thread logic that turns
// the summed reduction quantities of one step into the next scalars (alpha, omega,
// beta, Givens rotations, exit decisions).  Scalars never leave the device.
#pragma once
#include <math.h>

#include "internal.h"

namespace asam {

// Scalar logic of one finalize step; S = the summed reduction quantities.
__device__ inline void fin_logic(int kind, const double* S, KState* st, int k, int i, int nq,
                                 double* out) {
  switch (kind) {
    case FIN_SUMS:
      for (int q = 0; q < 4; ++q) st->sum[q] = S[q];
      break;
    case FIN_DOT:
      out[0] = S[0];
      break;
    case FIN_BICG_INIT: {
      // rNorm0 = ||r||; iteration 1 computes rho = rTilde.r = ||r||^2
      st->rnorm0 = sqrt(S[0]);
      st->rnorm = st->rnorm0;
      st->rho = S[0];
      st->iter = 1;
      st->done = 0; st->half = 0; st->pad = 0;
      st->status = 1;
      if (st->rho == 0.0) {  // reference: MaxIter = Iter, Tol = 0
        st->result = 1; st->rnorm = 0.0; st->status = 2; st->done = 1; st->pad = 1;
      } else if (!isfinite(st->rho)) {  // non-finite b or x0: stop, ||r|| stays non-finite
        st->result = 2; st->status = 1; st->done = 1; st->pad = 1;
      }
      break;
    }
    case FIN_BICG_ALPHA:
      if (st->done) break;
      st->alpha = st->rho / S[1];
      break;
    case FIN_BICG_HALF: {
      if (st->done) break;
      double rn = sqrt(S[0]);
      st->rnorm = rn;
      if (rn <= st->tol * st->rnorm0) {
        st->half = 1; st->pad = 1; st->result = -st->iter; st->status = 0;
      }
      break;
    }
    case FIN_BICG_OMEGA:
      if (st->done || st->pad) break;
      st->omega = S[1] / S[0];
      break;
    case FIN_BICG_HALF_OMEGA: {
      // SpMatrix4CellPar_Mod.f90:306-321: the half-step test on |s| comes first; the
      // t = MatVec(rHat) products only matter when it fails.
      if (st->done) break;
      double rn = sqrt(S[2]);
      st->rnorm = rn;
      if (rn <= st->tol * st->rnorm0) {
        st->half = 1; st->pad = 1; st->result = -st->iter; st->status = 0;
      } else {
        st->omega = S[1] / S[0];
      }
      break;
    }
    case FIN_BICG_END: {
      if (st->done) break;
      if (st->half) { st->done = 1; st->pad = 1; break; }
      double rn = sqrt(S[0]);
      st->rnorm = rn;
      if (rn <= st->tol * st->rnorm0) {
        st->result = st->iter; st->status = 0; st->done = 1;
      } else if (st->iter == st->maxiter || !isfinite(rn)) {
        // cap reached, or a non-finite iterate (rnorm keeps it for the caller)
        st->result = 2 * st->iter; st->status = 1; st->done = 1;
      } else {
        st->rho_old = st->rho;
        st->rho = S[1];
        if (st->rho == 0.0) {
          st->result = st->iter + 1; st->rnorm = 0.0; st->status = 2; st->done = 1;
        } else {
          st->beta = (st->rho / st->rho_old) * (st->alpha / st->omega);
          st->iter += 1;
        }
      }
      st->pad = st->done;
      break;
    }
    case FIN_GM_INIT: {
      // start of a restart cycle: beta = ||b - Op x||
      if (st->done) { st->pad = 1; st->cycle_done = 1; break; }
      double beta = sqrt(S[0]);
      if (st->cycles == 0) st->grho0 = beta;
      st->gbeta = beta;
      if (beta <= 0.0) {
        st->done = 1; st->pad = 1; st->cycle_done = 1; st->grho = 0.0; st->gstatus = 0;
        break;
      }
      if (!isfinite(beta)) {  // non-finite residual: no cycle, grho reports it
        st->done = 1; st->pad = 1; st->cycle_done = 1; st->grho = beta; st->gstatus = 1;
        break;
      }
      st->vscale[0] = 1.0 / beta;
      st->bhat[0] = beta * 1.0;
      st->nR = st->m;
      st->cycle_done = 0;
      st->pad = 0;
      break;
    }
    case FIN_GM_H:
      if (st->cycle_done) break;
      st->H[k * kMaxM + i] = (nq == 2) ? S[1] * st->vscale[0] : S[0];
      break;
    case FIN_GM_HV:
      // classical Gram-Schmidt: H[k + q][i] for the nq projections of one pass
      if (st->cycle_done) break;
      for (int q = 0; q < nq; ++q) st->H[(k + q) * kMaxM + i] = S[q];
      break;
    case FIN_GM_LAST: {
      if (st->cycle_done) break;
      double* H = st->H;
      const int M = kMaxM;
      H[(i + 1) * M + i] = sqrt(S[0]);
      st->vscale[i + 1] = 1.0 / H[(i + 1) * M + i];
      for (int q = 1; q <= i; ++q) {
        double temp = st->cs[q - 1] * H[(q - 1) * M + i] + st->sn[q - 1] * H[q * M + i];
        H[q * M + i] = -st->sn[q - 1] * H[(q - 1) * M + i] + st->cs[q - 1] * H[q * M + i];
        H[(q - 1) * M + i] = temp;
      }
      double delta =
          sqrt(H[i * M + i] * H[i * M + i] + H[(i + 1) * M + i] * H[(i + 1) * M + i]);
      st->cs[i] = H[i * M + i] / delta;
      st->sn[i] = H[(i + 1) * M + i] / delta;
      H[i * M + i] = st->cs[i] * H[i * M + i] + st->sn[i] * H[(i + 1) * M + i];
      st->bhat[i + 1] = -st->sn[i] * st->bhat[i];
      st->bhat[i] = st->cs[i] * st->bhat[i];
      st->grho = fabs(st->bhat[i + 1]);
      st->total += 1;
      if (st->grho <= st->gtol * st->grho0) {
        st->nR = i + 1;
        st->cycle_done = 1;
      } else if (!isfinite(st->grho)) {  // non-finite step: end the cycle and the solve
        st->nR = i + 1;
        st->cycle_done = 1;
        st->done = 1;
      } else {
        st->nR = st->m;
        if (i == st->m - 1) st->cycle_done = 1;
      }
      break;
    }
    case FIN_MIXED_OUTER: {
      // FP64 residual of the refinement step; sets up the FP32 inner BiCGStab on A d = r.
      if (st->outer_done) break;
      if (st->outer > 0) {  // inner iterations of the previous refinement step
        int res = st->result, it;
        if (st->done) it = res < 0 ? -res : (st->status == 1 ? res / 2 : res);
        else it = st->iter - 1;  // inner solve cut short by a fixed schedule
        st->inner_total += it;
        if (st->outer <= 16) st->hist[st->outer - 1] = it;
      }
      // x after this step's update x += d is in buffer (outer & 1) of the pair
      st->xpar = st->outer & 1;
      double r = sqrt(S[0]);
      if (st->outer == 0) st->outer_r0 = r;
      st->outer_r = r;
      if (r <= st->outer_tol * st->outer_r0 || r == 0.0 || !isfinite(r)) {
        // converged, or a non-finite residual (outer_r reports it; no further step)
        st->outer_done = 1; st->gstatus = isfinite(r) ? 0 : 1; st->done = 1; st->pad = 1;
        break;
      }
      // Ask the inner solve for inner_tol, but no more than the outer target needs.
      double need = 0.5 * st->outer_tol * st->outer_r0 / r;
      double itol = st->inner_tol;
      if (need > itol) itol = need;
      if (itol > 0.5) itol = 0.5;
      st->tol = itol;
      st->rnorm0 = r;
      st->rnorm = r;
      st->rho = S[0];  // rTilde = r
      st->iter = 1;
      st->done = 0; st->half = 0; st->pad = 0; st->status = 1;
      st->outer += 1;
      break;
    }
    case FIN_MIXED_GM: {
      // FP64 residual of the refinement step; sets up one FP32 (F)GMRES cycle on A d = r
      // from d = 0 (the restart of the reference's cycle loop is the next refinement step)
      if (st->outer_done) break;
      if (st->outer > 0) {  // Arnoldi steps of the previous step's cycle
        st->inner_total += st->total;
        if (st->outer <= 16) st->hist[st->outer - 1] = st->total;
      }
      st->xpar = st->outer & 1;
      const double r = sqrt(S[0]);
      if (st->outer == 0) st->outer_r0 = r;
      st->outer_r = r;
      if (r <= st->outer_tol * st->outer_r0 || r == 0.0 || !isfinite(r)) {
        st->outer_done = 1; st->gstatus = isfinite(r) ? 0 : 1; st->done = 1; st->pad = 1;
        st->cycle_done = 1;
        break;
      }
      double need = 0.5 * st->outer_tol * st->outer_r0 / r;
      double itol = st->inner_tol;
      if (need > itol) itol = need;
      if (itol > 0.5) itol = 0.5;
      st->gtol = itol;
      st->total = 0; st->cycles = 0;
      st->grho0 = r; st->gbeta = r; st->grho = r;
      st->vscale[0] = 1.0 / r;
      st->bhat[0] = r;
      st->nR = st->m;
      st->done = 0; st->pad = 0; st->cycle_done = 0;
      st->outer += 1;
      break;
    }
    default:
      break;
  }
}

}  // namespace asam
