// setup.hip -- operator set-up kernels (face weights, coupling coefficients, vertical-line
// LU factors, FP32 copies), the theta sensitivity of the operator, and the copy
// calibration kernel.
#include "kcommon.h"

namespace asam {

// ===================================================================== set-up
// WeightDiagCompute 'Density' (JacAccGrav_Mod.f90:4254-4270): theta = rho*theta / rho, and
// DUT = dp/d(rho*theta) of moist air with vapour rhov and condensate rhol + rhor (each
// pointer may be null: none of it):
//   RhoD = Rho - RhoV - RhoL + Eps, Rm = Rd + Rv RhoV/RhoD,
//   Cpml = Cpd + Cpv RhoV/RhoD + Cpl RhoL/RhoD, Kappa = Rm/Cpml,
//   DUT = T/(rho*theta + Eps) (RhoD Rd + RhoV Rv)/(1 - Kappa).
// Dry air gives Rm = Rd, Cpml = Cpd.
__global__ void k_weights(long N, const double* rho, const double* rth, const double* T,
                          const double* rhov, const double* rhol, const double* rhor,
                          double* theta, double* dut) {
  long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (c >= N) return;
  theta[c] = rth[c] / (rho[c] + kEps);
  const double rv = rhov ? rhov[c] : 0.0;
  const double rl = (rhol ? rhol[c] : 0.0) + (rhor ? rhor[c] : 0.0);
  const double rhoD = rho[c] - rv - rl + kEps;
  const double Rm = kRd + kRv * rv / rhoD;
  const double Cpml = kCpd + kCpv * rv / rhoD + kCpl * rl / rhoD;
  const double kappa = Rm / Cpml;
  dut[c] = T[c] / (rth[c] + kEps) * (rhoD * kRd + rv * kRv) / (1.0 - kappa);
}

void launch_weights(hipStream_t s, const Geom& g, const double* rho, const double* rth,
                    const double* T, const double* rhov, const double* rhol,
                    const double* rhor, double* theta, double* dut) {
  long N = g.N;
  hipLaunchKernelGGL(k_weights, dim3((N + 255) / 256), dim3(256), 0, s, N, rho, rth, T, rhov,
                     rhol, rhor, theta, dut);
}

// Volume-weighted face theta (CellToFaceRhoWeight); wall faces copy the adjacent cell.
__global__ void k_face_theta(int nx, int ny, int nz, int dir, const double* vol,
                             const double* th, double* out) {
  const int ex = nx + (dir == 0), ey = ny + (dir == 1), ez = nz + (dir == 2);
  long f = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long nf = (long)ex * ey * ez;
  if (f >= nf) return;
  int i = f % ex, j = (f / ex) % ey, k = f / ((long)ex * ey);
  int m = dir == 0 ? i : (dir == 1 ? j : k);
  int mmax = dir == 0 ? nx : (dir == 1 ? ny : nz);
  long sy = nx, sz = (long)nx * ny;
  long st = dir == 0 ? 1 : (dir == 1 ? sy : sz);
  long cb = i + sy * j + sz * k;
  // cell "right" of face m is index m, "left" is m-1 (along dir)
  if (m == 0) {
    out[f] = th[cb];
  } else if (m == mmax) {
    out[f] = th[cb - st];
  } else {
    long cr = cb, cl = cb - st;
    out[f] = (th[cr] * vol[cr] + th[cl] * vol[cl]) / (vol[cr] + vol[cl] + kEps);
  }
}

void launch_face_theta(hipStream_t s, const Geom& g, const double* vol, const double* theta,
                       double* thu, double* thv, double* thw) {
  long nfu = (long)(g.nx + 1) * g.ny * g.nz, nfv = (long)g.nx * (g.ny + 1) * g.nz,
       nfw = (long)g.nx * g.ny * (g.nz + 1);
  hipLaunchKernelGGL(k_face_theta, dim3((nfu + 255) / 256), dim3(256), 0, s, g.nx, g.ny, g.nz,
                     0, vol, theta, thu);
  hipLaunchKernelGGL(k_face_theta, dim3((nfv + 255) / 256), dim3(256), 0, s, g.nx, g.ny, g.nz,
                     1, vol, theta, thv);
  hipLaunchKernelGGL(k_face_theta, dim3((nfw + 255) / 256), dim3(256), 0, s, g.nx, g.ny, g.nz,
                     2, vol, theta, thw);
}

struct CoefArgs {
  int nx, ny, nz, np1, grad_full;
  double bd2, G;  // (beta0 dtP)^2, GravComp*(beta0 dtP)^2
  const double *FU, *FV, *FW, *vol, *dx, *dy, *dz, *thu, *thv, *thw;
  Coef<double> c;
};

// Coupling of the face between cells a and b (b the "+" neighbour) of one direction.
// MultiTriT:  k = (b0 dt)^2 2 F th DUU / (da+db)                 (ComputeFull_1D :78-83)
//             k = (b0 dt)^2 2 F th DUU F / (Va+Vb+Eps)           (ComputeDual_1D :48-53)
// MultiTriTB: k' = (b0 dt)^2 2 F DUU FG / (Va+Vb+Eps), FG per ReadWeights_Mod.f90:1376-1462
__device__ __forceinline__ void face_coef(const CoefArgs& a, double F, double th, double Va,
                                          double Vb, double da, double db, double& k,
                                          double& t) {
  const double DUU = F / (F + kEps);
  if (a.np1 == 1) {
    if (a.grad_full) k = a.bd2 * 2.0 * F * th * DUU / ((da + db) * 1.0 * 1.0);
    else k = a.bd2 * 2.0 * F * th * DUU * F / (Va + Vb + kEps);
    t = th;
  } else {
    const double FG = a.grad_full ? (Va + Vb) / (1.0 * (db + da + kEps)) * F / (F + kEps) : F;
    k = a.bd2 * 2.0 * F * DUU * FG / (Va + Vb + kEps);
    t = th;
  }
}

__global__ void k_coef(CoefArgs a) {
  long N = (long)a.nx * a.ny * a.nz;
  long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (c >= N) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  int i = c % nx, j = (c / nx) % ny, k = c / ((long)nx * ny);
  const long sy = nx, sz = (long)nx * ny;
  double kk, tt;
  // x-face i+1
  if (i + 1 < nx) {
    long f = (i + 1) + (long)(nx + 1) * (j + (long)ny * k);
    face_coef(a, a.FU[f], a.thu[f], a.vol[c], a.vol[c + 1], a.dx[i], a.dx[i + 1], kk, tt);
  } else { kk = 0.0; tt = 0.0; }
  a.c.kx[c] = kk;
  if (a.np1 == 2) a.c.tx[c] = tt;
  if (j + 1 < ny) {
    long f = i + (long)nx * ((j + 1) + (long)(ny + 1) * k);
    face_coef(a, a.FV[f], a.thv[f], a.vol[c], a.vol[c + sy], a.dy[j], a.dy[j + 1], kk, tt);
  } else { kk = 0.0; tt = 0.0; }
  a.c.ky[c] = kk;
  if (a.np1 == 2) a.c.ty[c] = tt;
  double gz = 0.0;
  if (k + 1 < nz) {
    long f = i + (long)nx * (j + (long)ny * (k + 1));
    face_coef(a, a.FW[f], a.thw[f], a.vol[c], a.vol[c + sz], a.dz[k], a.dz[k + 1], kk, tt);
    gz = a.G * a.FW[f] / (a.vol[c] + a.vol[c + sz] + kEps);
  } else { kk = 0.0; tt = 0.0; }
  a.c.kz[c] = kk;
  if (a.np1 == 2) { a.c.tz[c] = tt; a.c.gz[c] = gz; }
  a.c.vol[c] = a.vol[c];
}

void launch_coef(hipStream_t s, const Geom& g, int grad_full, double beta0_dt, double grav,
                 const double* FU, const double* FV, const double* FW, const double* vol,
                 const double* dx, const double* dy, const double* dz, const double* thu,
                 const double* thv, const double* thw, Coef<double>& c) {
  CoefArgs a;
  a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.np1 = g.np1; a.grad_full = grad_full;
  a.bd2 = beta0_dt * beta0_dt;
  a.G = grav * a.bd2;
  a.FU = FU; a.FV = FV; a.FW = FW; a.vol = vol; a.dx = dx; a.dy = dy; a.dz = dz;
  a.thu = thu; a.thv = thv; a.thw = thw; a.c = c;
  hipLaunchKernelGGL(k_coef, dim3((g.N + 255) / 256), dim3(256), 0, s, a);
}

// ----------------------------------------------------------- line factorisation
// LU of each vertical line (Solve_1D, TriDiagT_Mod.f90:135-155; block form with
// Div_22/Mult_22, TriDiagTB_Mod.f90:143-163):  t_0 = d_0,  w_k-1 = t_k-1^-1 u_k-1,
// t_k = d_k - l_k-1 w_k-1.  Stored per level k: pl = l_k-1 (0 at k=0),
// pt = t_k^-1 (0 for a singular t: a dead cell), pw = w_k (0 at the top).
__device__ __forceinline__ void inv22(const double* m, double* r) {
  double det = m[0] * m[3] - m[2] * m[1];
  if (det == 0.0) { r[0] = r[1] = r[2] = r[3] = 0.0; return; }
  r[0] = m[3] / det; r[1] = -m[1] / det; r[2] = -m[2] / det; r[3] = m[0] / det;
}
__device__ __forceinline__ void mul22(const double* a, const double* b, double* r) {
  double t0 = a[0] * b[0] + a[2] * b[1];
  double t1 = a[1] * b[0] + a[3] * b[1];
  double t2 = a[0] * b[2] + a[2] * b[3];
  double t3 = a[1] * b[2] + a[3] * b[3];
  r[0] = t0; r[1] = t1; r[2] = t2; r[3] = t3;
}

__global__ void k_factor(Geom g, Coef<double> c) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  long col = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (col >= (long)nx * ny) return;
  const int i = col % nx, j = col / nx;
  const long sy = nx, sz = (long)nx * ny;
  if (g.np1 == 1) {
    double t = 0.0, wprev = 0.0;
    for (int k = 0; k < nz; ++k) {
      long cc = col + sz * k;
      const double dut = c.dut[cc];
      double d = g.fa * c.vol[cc];
      if (i > 0) d += c.kx[cc - 1] * dut;
      if (i + 1 < nx) d += c.kx[cc] * dut;
      if (j > 0) d += c.ky[cc - sy] * dut;
      if (j + 1 < ny) d += c.ky[cc] * dut;
      if (k > 0) d += c.kz[cc - sz] * dut;
      if (k + 1 < nz) d += c.kz[cc] * dut;
      double l = (k > 0) ? -c.kz[cc - sz] * c.dut[cc - sz] : 0.0;  // row k, col k-1
      t = (k == 0) ? d : d - l * wprev;
      double tinv = (t == 0.0) ? 0.0 : 1.0 / t;
      double w = 0.0;
      if (k + 1 < nz) {
        double u = -c.kz[cc] * c.dut[cc + sz];  // row k, col k+1
        w = (t == 0.0) ? 0.0 : u / t;
      }
      const long fi = fac_index(g, i, j, k);
      c.pl[fi] = l;
      c.pt[fi] = tinv;
      c.pw[fi] = w;
      wprev = w;
    }
  } else {
    double W[4] = {0, 0, 0, 0};
    for (int k = 0; k < nz; ++k) {
      long cc = col + sz * k;
      const double dut = c.dut[cc], V = c.vol[cc];
      double d[4];
      d[0] = g.fa * V; d[1] = 0.0; d[2] = 0.0; d[3] = V;
      auto hface = [&](double kk, double th) { d[0] += th * kk * dut; d[1] += kk * dut; };
      if (i > 0) hface(c.kx[cc - 1], c.tx[cc - 1]);
      if (i + 1 < nx) hface(c.kx[cc], c.tx[cc]);
      if (j > 0) hface(c.ky[cc - sy], c.ty[cc - sy]);
      if (j + 1 < ny) hface(c.ky[cc], c.ty[cc]);
      double L[4] = {0, 0, 0, 0};
      double lz_tz = 0.0;  // tz of the face below (L's first row is tz times its second)
      if (k > 0) {
        const long m = cc - sz;
        const double kz = c.kz[m], tz = c.tz[m], gz = c.gz[m];
        lz_tz = tz;
        hface(kz, tz);
        d[2] += tz * gz * V;   // +th g~ V_c   (d(1,2) of the upper cell)
        d[3] += gz * V;
        L[0] = -tz * kz * c.dut[m];
        L[1] = -kz * c.dut[m];
        L[2] = tz * gz * c.vol[m];
        L[3] = gz * c.vol[m];
      }
      double U[4] = {0, 0, 0, 0};
      if (k + 1 < nz) {
        const double kz = c.kz[cc], tz = c.tz[cc], gz = c.gz[cc];
        hface(kz, tz);
        d[2] -= tz * gz * V;
        d[3] -= gz * V;
        const long p = cc + sz;
        U[0] = -tz * kz * c.dut[p];
        U[1] = -kz * c.dut[p];
        U[2] = -tz * gz * c.vol[p];
        U[3] = -gz * c.vol[p];
      }
      double t[4];
      if (k == 0) {
        t[0] = d[0]; t[1] = d[1]; t[2] = d[2]; t[3] = d[3];
      } else {
        double LW[4];
        mul22(L, W, LW);
        for (int q = 0; q < 4; ++q) t[q] = d[q] - LW[q];
      }
      double T[4];
      inv22(t, T);
      double Wn[4] = {0, 0, 0, 0};
      if (k + 1 < nz) mul22(T, U, Wn);
      const long fi = fac_index(g, i, j, k);
      for (int q = 0; q < 4; ++q) {
        c.pt[4 * fi + q] = T[q];
        c.pw[4 * fi + q] = Wn[q];
        W[q] = Wn[q];
      }
      // L = (tz, 1)^T (L[1], L[3]): the sweeps use the three words
      c.pl[3 * fi + 0] = lz_tz;
      c.pl[3 * fi + 1] = L[1];
      c.pl[3 * fi + 2] = L[3];
    }
  }
}

void launch_factor(hipStream_t s, const Geom& g, Coef<double>& c) {
  long ncol = (long)g.nx * g.ny;
  hipLaunchKernelGGL(k_factor, dim3((ncol + 63) / 64), dim3(64), 0, s, g, c);
}

__global__ void k_cvt(long n, const double* a, float* b) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void k_pack_z(long n, Coef<float> c) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  reinterpret_cast<float4*>(c.pkz)[i] = make_float4(c.kz[i], c.tz[i], c.gz[i], c.vol[i]);
}

void launch_convert(hipStream_t s, const Geom& g, const Coef<double>& c64, Coef<float>& c32) {
  const long N = g.N, P = (long)g.np1 * g.np1 * fac_cells(g);
  const long PL = (long)lower_words(g.np1) * fac_cells(g);
  auto cv = [&](const double* a, float* b, long n) {
    if (a && b) hipLaunchKernelGGL(k_cvt, dim3((n + 255) / 256), dim3(256), 0, s, n, a, b);
  };
  cv(c64.kx, c32.kx, N); cv(c64.ky, c32.ky, N); cv(c64.kz, c32.kz, N);
  cv(c64.tx, c32.tx, N); cv(c64.ty, c32.ty, N); cv(c64.tz, c32.tz, N);
  cv(c64.gz, c32.gz, N); cv(c64.vol, c32.vol, N); cv(c64.dut, c32.dut, N);
  cv(c64.pl, c32.pl, PL); cv(c64.pt, c32.pt, P); cv(c64.pw, c32.pw, P);
  if (c32.pkz) hipLaunchKernelGGL(k_pack_z, dim3((N + 255) / 256), dim3(256), 0, s, N, c32);
}

// Dead cells of the residual form: the apply's own test (march_tile), same operand order.
__global__ void k_dead_mask(int nx, int ny, int nz, const double* kx, const double* ky,
                            const double* kz, const double* vol, unsigned char* dead) {
  const long N = (long)nx * ny * nz, sy = nx, sz = (long)nx * ny;
  const long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (c >= N) return;
  const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / sz);
  const long mW = i > 0 ? c - 1 : c, mS = j > 0 ? c - sy : c;
  const double kzm = k > 0 ? kz[c - sz] : 0.0;
  dead[c] = (vol[c] == 0.0 && kx[mW] + kx[c] + ky[mS] + ky[c] + kzm + kz[c] == 0.0) ? 1 : 0;
}

void launch_dead_mask(hipStream_t s, const Geom& g, const Coef<double>& c64, unsigned char* dead) {
  hipLaunchKernelGGL(k_dead_mask, dim3((g.N + 255) / 256), dim3(256), 0, s, g.nx, g.ny, g.nz,
                     c64.kx, c64.ky, c64.kz, c64.vol, dead);
}

// ===================================================================== theta VJP
// d(mu^T A x)/d theta_c through the face theta th_f = (th_a V_a + th_b V_b)/(V_a+V_b+Eps).
// MultiTriT: dS/dth_f = (k_f/th_f) (q_a - q_b)(mu_a - mu_b)
// MultiTriTB: dS/dth_f = k'_f (q_a - q_b)(mu1_a - mu1_b) + [z] B_f (mu1_b - mu1_a)
__global__ void k_theta_vjp(Geom g, Coef<double> C, const double* thu, const double* thv,
                            const double* thw, double* grad, const double* mu,
                            const double* x) {
  const int nx = g.nx, ny = g.ny, nz = g.nz, NP = g.np1;
  long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (c >= g.N) return;
  const int i = c % nx, j = (c / nx) % ny, k = c / ((long)nx * ny);
  const long sy = nx, sz = (long)nx * ny;
  auto X1 = [&](long m) { return NP == 1 ? x[m] : x[2 * m]; };
  auto X2 = [&](long m) { return x[2 * m + 1]; };
  auto M1 = [&](long m) { return NP == 1 ? mu[m] : mu[2 * m]; };
  double gsum = 0.0;
  // face between a and b = a + e (b "+"), f stored at a; th_face from full face arrays.
  auto face = [&](long a, long b, double kk, double thf, double gzz, bool vert) {
    double qa = C.dut[a] * X1(a), qb = C.dut[b] * X1(b);
    double dS;
    if (NP == 1) {
      dS = (thf != 0.0 ? kk / thf : 0.0) * (qa - qb) * (M1(a) - M1(b));
    } else {
      dS = kk * (qa - qb) * (M1(a) - M1(b));
      if (vert) {
        double B = gzz * (C.vol[a] * X2(a) + C.vol[b] * X2(b));
        dS += B * (M1(b) - M1(a));
      }
    }
    double w = C.vol[c] / (C.vol[a] + C.vol[b] + kEps);
    gsum += dS * w;
  };
  if (i > 0) {
    long f = i + (long)(nx + 1) * (j + (long)ny * k);
    face(c - 1, c, C.kx[c - 1], thu[f], 0.0, false);
  }
  if (i + 1 < nx) {
    long f = (i + 1) + (long)(nx + 1) * (j + (long)ny * k);
    face(c, c + 1, C.kx[c], thu[f], 0.0, false);
  }
  if (j > 0) {
    long f = i + (long)nx * (j + (long)(ny + 1) * k);
    face(c - sy, c, C.ky[c - sy], thv[f], 0.0, false);
  }
  if (j + 1 < ny) {
    long f = i + (long)nx * ((j + 1) + (long)(ny + 1) * k);
    face(c, c + sy, C.ky[c], thv[f], 0.0, false);
  }
  if (k > 0) {
    long f = i + (long)nx * (j + (long)ny * k);
    face(c - sz, c, C.kz[c - sz], thw[f], NP == 2 ? C.gz[c - sz] : 0.0, true);
  }
  if (k + 1 < nz) {
    long f = i + (long)nx * (j + (long)ny * (k + 1));
    face(c, c + sz, C.kz[c], thw[f], NP == 2 ? C.gz[c] : 0.0, true);
  }
  grad[c] += gsum;
}

void launch_theta_vjp(hipStream_t s, const Geom& g, const Coef<double>& c,
                      const double* thfaces, double* grad, const double* mu, const double* x) {
  const double* thu = thfaces;
  const double* thv = thu + (long)(g.nx + 1) * g.ny * g.nz;
  const double* thw = thv + (long)g.nx * (g.ny + 1) * g.nz;
  hipLaunchKernelGGL(k_theta_vjp, dim3((g.N + 255) / 256), dim3(256), 0, s, g, c, thu, thv,
                     thw, grad, mu, x);
}

// Streaming copy (16-byte vectors, one per thread: the fastest copy shape measured on
// gfx950) used to calibrate achievable HBM bandwidth and the profiler's FETCH_SIZE /
// WRITE_SIZE counters against a kernel of known traffic.
__global__ __launch_bounds__(256) void k_copy16(long n16, const float4* __restrict__ a,
                                                float4* __restrict__ b) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < n16) b[e] = a[e];
}

void launch_copy16(hipStream_t s, long n16, const void* a, void* b) {
  const long nb = (n16 + 255) / 256;
  if (nb < 1 || nb > 0x7fffffffL) return;
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)nb), dim3(256), 0, s, n16, (const float4*)a,
                     (float4*)b);
}

}  // namespace asam
