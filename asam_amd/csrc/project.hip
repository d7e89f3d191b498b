// project.hip -- the two steps either side of the stage solve, on the GPU (gfx950):
// the right-hand side from the predicted face velocities and the stage increments (the
// reference's Rhs, MATRIX/Rhs_Mod.f90:110-157) and, after the solve, the face-velocity and
// increment update (UpdateVelocity, :589-952).  One closed block, dry 'Density' weights
// (DTT = 1, DUR = 0), Cartesian metric.
//
// Three flat kernels, one thread per cell in x-fastest order (the access order that keeps
// the most HBM streams on open pages): the divergence, the correction of each cell's
// east/north/top interior faces, and the cell update from the corrected fluxes.  Every
// sum keeps the reference's term order.
#include "kcommon.h"

namespace asam {

struct ProjGeom {
  int nx, ny, nz, np1, grad_full;
  double bd, grav, fa;
  const double *FU, *FV, *FW;     // face areas, reference face shapes
  const double *thu, *thv, *thw;  // face theta
  const double *V, *dut, *dx, *dy, *dz;
};

__device__ __forceinline__ void cell_ijk(const ProjGeom& g, long c, int& i, int& j, int& k) {
  i = (int)(c % g.nx);
  const long r = c / g.nx;
  j = (int)(r % g.ny);
  k = (int)(r / g.ny);
}

// b1 = bd * DivergenceS - V incr_rt (fa >= 1); TB: b2 = bd * Divergence - V incr_r
__global__ __launch_bounds__(256) void k_rhs(ProjGeom g, double* __restrict__ b,
                                             const double* __restrict__ U,
                                             const double* __restrict__ Vv,
                                             const double* __restrict__ W,
                                             const double* __restrict__ incr_rt,
                                             const double* __restrict__ incr_r, long N) {
  const long c = blockIdx.x * 256L + threadIdx.x;
  if (c >= N) return;
  int i, j, k;
  cell_ijk(g, c, i, j, k);
  const long nx = g.nx, ny = g.ny;
  const long uw = i + (nx + 1) * (j + ny * k), ue = uw + 1;
  const long vs = i + nx * (j + (ny + 1) * k), vn = vs + nx;
  const long wb = c, wt = c + nx * ny;
  double s = -g.FU[uw] * g.thu[uw] * U[uw];
  s += g.FU[ue] * g.thu[ue] * U[ue];
  s += -g.FV[vs] * g.thv[vs] * Vv[vs];
  s += g.FV[vn] * g.thv[vn] * Vv[vn];
  s += -g.FW[wb] * g.thw[wb] * W[wb];
  s += g.FW[wt] * g.thw[wt] * W[wt];
  double b1 = g.bd * s;
  if (g.fa >= 1.0) b1 -= g.V[c] * incr_rt[c];
  if (g.np1 == 1) {
    b[c] = b1;
  } else {
    double r = -g.FU[uw] * U[uw];
    r += g.FU[ue] * U[ue];
    r += -g.FV[vs] * Vv[vs];
    r += g.FV[vn] * Vv[vn];
    r += -g.FW[wb] * W[wb];
    r += g.FW[wt] * W[wt];
    st2(b, c, b1, g.bd * r - g.V[c] * incr_r[c]);
  }
}

// Interior faces east/north/top of cell c: u -= grad(DUT x1) (GradFull or Dual form);
// TB: the top face first takes the buoyancy correction with x2.
template <int NP>
__global__ __launch_bounds__(256) void k_update_faces(ProjGeom g, const double* __restrict__ x,
                                                      double* __restrict__ U,
                                                      double* __restrict__ Vv,
                                                      double* __restrict__ W, long N) {
  const long c = blockIdx.x * 256L + threadIdx.x;
  if (c >= N) return;
  int i, j, k;
  cell_ijk(g, c, i, j, k);
  const long nx = g.nx, ny = g.ny, sy = nx, sz = nx * ny;
  const double bd = g.bd;
  const double q0 = x[NP * c] * g.dut[c];
  double p2c = 0.0;
  if constexpr (NP == 2) p2c = x[2 * c + 1] * 0.0;  // x2 * DUR
  auto dpn = [&](long n) {
    double dp = x[NP * n] * g.dut[n] - q0;
    if constexpr (NP == 2) dp = dp + (x[2 * n + 1] * 0.0 - p2c);
    return dp;
  };
  if (i + 1 < g.nx) {
    const long f = (i + 1) + (nx + 1) * (j + ny * k), n = c + 1;
    const double F = g.FU[f], duu = F / (F + kEps), dp = dpn(n);
    const double t = g.grad_full ? 2.0 * bd * duu * dp / (g.dx[i + 1] + g.dx[i])
                                 : 2.0 * bd * F * duu * dp / (g.V[n] + g.V[c] + kEps);
    U[f] = U[f] - t;
  }
  if (j + 1 < g.ny) {
    const long f = i + nx * ((j + 1) + (ny + 1) * k), n = c + sy;
    const double F = g.FV[f], duv = F / (F + kEps), dp = dpn(n);
    const double t = g.grad_full ? 2.0 * bd * duv * dp / (g.dy[j + 1] + g.dy[j])
                                 : 2.0 * bd * F * duv * dp / (g.V[n] + g.V[c] + kEps);
    Vv[f] = Vv[f] - t;
  }
  if (k + 1 < g.nz) {
    const long f = c + sz, n = c + sz;
    const double F = g.FW[f], duw = F / (F + kEps);
    double w = W[f];
    if constexpr (NP == 2) {
      const double tg = bd * g.grav * (g.V[c] * x[2 * c + 1] + g.V[n] * x[2 * n + 1]) /
                        (g.V[n] + g.V[c] + kEps);
      w = w - tg;
    }
    const double dp = dpn(n);
    if (g.grad_full) {
      const double t = 2.0 * bd * duw * dp / (g.dz[k + 1] + g.dz[k]);
      w = NP == 1 ? w - F / (F + kEps) * t : w - t;
    } else {
      w = w - 2.0 * bd * F * duw * dp / (g.V[n] + g.V[c] + kEps);
    }
    W[f] = w;
  }
}

// Increments from the corrected face fluxes; x1 *= DUT, x2 *= DUR (= 0).
template <int NP>
__global__ __launch_bounds__(256) void k_update_cells(ProjGeom g, double* __restrict__ x,
                                                      const double* __restrict__ U,
                                                      const double* __restrict__ Vv,
                                                      const double* __restrict__ W,
                                                      double* __restrict__ incr_rt,
                                                      double* __restrict__ incr_r, long N) {
  const long c = blockIdx.x * 256L + threadIdx.x;
  if (c >= N) return;
  int i, j, k;
  cell_ijk(g, c, i, j, k);
  const long nx = g.nx, ny = g.ny;
  const long uw = i + (nx + 1) * (j + ny * k), ue = uw + 1;
  const long vs = i + nx * (j + (ny + 1) * k), vn = vs + nx;
  const long wb = c, wt = c + nx * ny;
  double s = g.FU[ue] * g.thu[ue] * U[ue];
  s += -g.FU[uw] * g.thu[uw] * U[uw];
  s += g.FV[vn] * g.thv[vn] * Vv[vn];
  s += -g.FV[vs] * g.thv[vs] * Vv[vs];
  s += g.FW[wt] * g.thw[wt] * W[wt];
  s += -g.FW[wb] * g.thw[wb] * W[wb];
  const double vc = g.V[c] + kEps;
  incr_rt[c] = incr_rt[c] - g.bd * s / vc;
  if (incr_r) {
    double r = g.FU[ue] * U[ue];
    r += -g.FU[uw] * U[uw];
    r += g.FV[vn] * Vv[vn];
    r += -g.FV[vs] * Vv[vs];
    r += g.FW[wt] * W[wt];
    r += -g.FW[wb] * W[wb];
    incr_r[c] = incr_r[c] - g.bd * r / vc;
  }
  if constexpr (NP == 1) {
    x[c] = x[c] * g.dut[c];
  } else {
    const double2 xv = *reinterpret_cast<const double2*>(x + 2 * c);
    st2(x, c, xv.x * g.dut[c], xv.y * 0.0);
  }
}

static ProjGeom proj_geom(const Geom& g, const ProjInputs& in) {
  ProjGeom p;
  p.nx = g.nx; p.ny = g.ny; p.nz = g.nz; p.np1 = g.np1; p.grad_full = in.grad_full;
  p.bd = in.beta0_dt; p.grav = in.grav; p.fa = g.fa;
  p.FU = in.FU; p.FV = in.FV; p.FW = in.FW;
  p.thu = in.thu; p.thv = in.thv; p.thw = in.thw;
  p.V = in.V; p.dut = in.dut; p.dx = in.dx; p.dy = in.dy; p.dz = in.dz;
  return p;
}

void launch_rhs(hipStream_t s, const Geom& g, const ProjInputs& in, double* b, const double* U,
                const double* V, const double* W, const double* incr_rt, const double* incr_r) {
  const ProjGeom p = proj_geom(g, in);
  const long nb = (g.N + 255) / 256;
  hipLaunchKernelGGL(k_rhs, dim3((unsigned)nb), dim3(256), 0, s, p, b, U, V, W, incr_rt, incr_r,
                     g.N);
}

void launch_update(hipStream_t s, const Geom& g, const ProjInputs& in, double* x, double* U,
                   double* V, double* W, double* incr_rt, double* incr_r) {
  const ProjGeom p = proj_geom(g, in);
  const long nb = (g.N + 255) / 256;
  if (g.np1 == 1) {
    hipLaunchKernelGGL(k_update_faces<1>, dim3((unsigned)nb), dim3(256), 0, s, p,
                       (const double*)x, U, V, W, g.N);
    hipLaunchKernelGGL(k_update_cells<1>, dim3((unsigned)nb), dim3(256), 0, s, p, x,
                       (const double*)U, (const double*)V, (const double*)W, incr_rt, incr_r,
                       g.N);
  } else {
    hipLaunchKernelGGL(k_update_faces<2>, dim3((unsigned)nb), dim3(256), 0, s, p,
                       (const double*)x, U, V, W, g.N);
    hipLaunchKernelGGL(k_update_cells<2>, dim3((unsigned)nb), dim3(256), 0, s, p, x,
                       (const double*)U, (const double*)V, (const double*)W, incr_rt, incr_r,
                       g.N);
  }
}

}  // namespace asam
