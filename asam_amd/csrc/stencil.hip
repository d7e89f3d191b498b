// stencil.hip -- set-up, stencil apply and vertical-line solve kernels (gfx950).
//
// Operator (one block, Neumann walls), per cell c with faces f to neighbours n:
//   MultiTriT : (A x)_c = fa V_c x_c + sum_f k_f (DUT_c x_c - DUT_n x_n)
//   MultiTriTB: row1 = fa V x1 + sum_f th_f k'_f (q_c - q_n) + vertical gravity * th_f
//               row2 =    V x2 + sum_f      k'_f (q_c - q_n) + vertical gravity
//               q = DUT x1; gravity flux of the z-face (k,k+1):
//               B = g~ (V_k x2_k + V_k+1 x2_k+1), -B into cell k, +B into cell k+1.
// This is the reference's nested tri-diagonal matrix (ComputeFull/Dual_1D..3D,
// TriDiagT_Mod.f90:32-900; Compute_1D..3D, TriDiagTB_Mod.f90:32-780) written per face:
// u = -k DUT_n, l = -k DUT_c, d += k DUT_c, and the 2x2 gravity entries
// (TriDiagTB_Mod.f90:56-81) with g~ = Grav (beta0 dtP)^2 F / (V_k + V_k+1 + Eps).
//
// Kernel shapes: apply = 2.5D march (64x4 threads over x-y, each thread walks a
// chunk of z keeping the column in registers; x-neighbours share L1 lines, y
// and z neighbours are re-read from L2/registers).  Line solve = one thread per
// vertical column, 64 columns per wavefront (coalesced along x at every level),
// forward-sweep results staged in LDS for the backward sweep.
#include "internal.h"

namespace asam {

// ===================================================================== helpers
template <int NQ>
__device__ __forceinline__ void block_reduce_store(double (&v)[NQ], double* partial, int blk,
                                                   int tid, int nthreads) {
  __shared__ double sh[NQ][16];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
  }
  const int lane = tid & 63, wid = tid >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) sh[q][wid] = v[q];
  }
  __syncthreads();
  if (tid == 0) {
    const int nw = (nthreads + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += sh[q][w];
      partial[q * kRedBlocks + blk] = s;
    }
  }
}

template <typename T> struct V2 { T a, b; };

template <typename T>
__device__ __forceinline__ V2<T> ld2(const T* p, long c) {
  if constexpr (sizeof(T) == 8) {
    double2 v = *reinterpret_cast<const double2*>(p + 2 * c);
    return {(T)v.x, (T)v.y};
  } else {
    float2 v = *reinterpret_cast<const float2*>(p + 2 * c);
    return {(T)v.x, (T)v.y};
  }
}
template <typename T, typename U>
__device__ __forceinline__ void st2(T* p, long c, U a, U b) {
  if constexpr (sizeof(T) == 8) {
    *reinterpret_cast<double2*>(p + 2 * c) = make_double2((double)a, (double)b);
  } else {
    *reinterpret_cast<float2*>(p + 2 * c) = make_float2((float)a, (float)b);
  }
}

// ===================================================================== set-up
__global__ void k_weights(long N, const double* rho, const double* rth, const double* T,
                          double* theta, double* dut) {
  long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (c >= N) return;
  theta[c] = rth[c] / (rho[c] + kEps);
  // WeightDiagCompute 'Density', dry air (RhoV = RhoL = 0): Rm = Rd, Cpml = Cpd.
  const double kappa = kRd / kCpd;
  const double rhoD = rho[c] - 0.0 - 0.0 + kEps;
  dut[c] = T[c] / (rth[c] + kEps) * (rhoD * kRd + 0.0) / (1.0 - kappa);
}

void launch_weights(hipStream_t s, const Geom& g, const double* rho, const double* rth,
                    const double* T, double* theta, double* dut) {
  long N = g.N;
  hipLaunchKernelGGL(k_weights, dim3((N + 255) / 256), dim3(256), 0, s, N, rho, rth, T, theta,
                     dut);
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
      c.pl[cc] = l;
      c.pt[cc] = tinv;
      c.pw[cc] = w;
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
      if (k > 0) {
        const long m = cc - sz;
        const double kz = c.kz[m], tz = c.tz[m], gz = c.gz[m];
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
      for (int q = 0; q < 4; ++q) {
        c.pl[4 * cc + q] = L[q];
        c.pt[4 * cc + q] = T[q];
        c.pw[4 * cc + q] = Wn[q];
        W[q] = Wn[q];
      }
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

void launch_convert(hipStream_t s, const Geom& g, const Coef<double>& c64, Coef<float>& c32) {
  const long N = g.N, P = (long)g.np1 * g.np1 * N;
  auto cv = [&](const double* a, float* b, long n) {
    if (a && b) hipLaunchKernelGGL(k_cvt, dim3((n + 255) / 256), dim3(256), 0, s, n, a, b);
  };
  cv(c64.kx, c32.kx, N); cv(c64.ky, c32.ky, N); cv(c64.kz, c32.kz, N);
  cv(c64.tx, c32.tx, N); cv(c64.ty, c32.ty, N); cv(c64.tz, c32.tz, N);
  cv(c64.gz, c32.gz, N); cv(c64.vol, c32.vol, N); cv(c64.dut, c32.dut, N);
  cv(c64.pl, c32.pl, P); cv(c64.pt, c32.pt, P); cv(c64.pw, c32.pw, P);
}

// ===================================================================== apply
template <typename CT, typename VX, typename VO, typename VU>
struct KApply {
  Coef<CT> c;
  int nx, ny, nz, kchunk;
  CT fa;
  const VX* x;
  const float* dx;
  double* xw;
  const double* b;
  VO* y;
  VO* y2;
  const VU* u;
  int dots;
  double* partial;
  const int* done;
};

int apply_blocks(const Geom& g) {
  int gx = (g.nx + 63) / 64, gy = (g.ny + 3) / 4;
  int cols = gx * gy;
  int nzc = (2048 + cols - 1) / cols;  // aim for ~2048 workgroups
  if (nzc < 1) nzc = 1;
  if (nzc > g.nz) nzc = g.nz;
  int kchunk = (g.nz + nzc - 1) / nzc;
  nzc = (g.nz + kchunk - 1) / kchunk;
  return cols * nzc;
}

template <int NP, int TR, typename CT, typename VX, typename VO, typename VU>
__global__ __launch_bounds__(256) void k_apply(KApply<CT, VX, VO, VU> a) {
  if (a.done && *a.done) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const long sy = nx, sz = (long)nx * ny;
  const int tid = threadIdx.x + 64 * threadIdx.y;
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
  const int k0 = blockIdx.z * a.kchunk, k1 = min(nz, k0 + a.kchunk);
  const int blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  double acc[2] = {0.0, 0.0};
  const Coef<CT>& C = a.c;
  const CT fa = a.fa;
  if (i < nx && j < ny && k0 < k1) {
    long c = i + sy * j + sz * k0;
    if constexpr (NP == 1) {
      auto LX = [&](long m) -> CT {
        CT v = (CT)a.x[m];
        if (a.dx) v += (CT)a.dx[m];
        return v;
      };
      CT xm = 0, dm = 0, kzm = 0;
      if (k0 > 0) { xm = LX(c - sz); dm = C.dut[c - sz]; kzm = C.kz[c - sz]; }
      CT x0 = LX(c), d0 = C.dut[c];
      for (int k = k0; k < k1; ++k, c += sz) {
        CT xp = 0, dp = 0;
        if (k + 1 < nz) { xp = LX(c + sz); dp = C.dut[c + sz]; }
        const CT kz0 = C.kz[c];
        CT s = fa * C.vol[c] * x0;
        if constexpr (TR == 0) {
          const CT q0 = d0 * x0;
          if (i > 0) s += C.kx[c - 1] * (q0 - C.dut[c - 1] * LX(c - 1));
          if (i + 1 < nx) s += C.kx[c] * (q0 - C.dut[c + 1] * LX(c + 1));
          if (j > 0) s += C.ky[c - sy] * (q0 - C.dut[c - sy] * LX(c - sy));
          if (j + 1 < ny) s += C.ky[c] * (q0 - C.dut[c + sy] * LX(c + sy));
          if (k > 0) s += kzm * (q0 - dm * xm);
          if (k + 1 < nz) s += kz0 * (q0 - dp * xp);
        } else {
          CT h = 0;
          if (i > 0) h += C.kx[c - 1] * (x0 - LX(c - 1));
          if (i + 1 < nx) h += C.kx[c] * (x0 - LX(c + 1));
          if (j > 0) h += C.ky[c - sy] * (x0 - LX(c - sy));
          if (j + 1 < ny) h += C.ky[c] * (x0 - LX(c + sy));
          if (k > 0) h += kzm * (x0 - xm);
          if (k + 1 < nz) h += kz0 * (x0 - xp);
          s += d0 * h;
        }
        // Op = -A x; residual form y = b - Op(x) = b + A x
        double yv = a.b ? a.b[c] + (double)s : -(double)s;
        if (a.xw) a.xw[c] = (double)x0;
        a.y[c] = (VO)yv;
        if (a.y2) a.y2[c] = (VO)yv;
        if (a.dots & DOT_YY) acc[0] += yv * yv;
        if (a.dots & DOT_YU) acc[1] += yv * (double)a.u[c];
        xm = x0; dm = d0; kzm = kz0; x0 = xp; d0 = dp;
      }
    } else {
      auto LX = [&](long m) -> V2<CT> {
        V2<VX> v = ld2(a.x, m);
        V2<CT> r{(CT)v.a, (CT)v.b};
        if (a.dx) { V2<float> e = ld2(a.dx, m); r.a += (CT)e.a; r.b += (CT)e.b; }
        return r;
      };
      V2<CT> xm{0, 0};
      CT dm = 0, vm = 0, kzm = 0, tzm = 0, gzm = 0;
      if (k0 > 0) {
        const long m = c - sz;
        xm = LX(m); dm = C.dut[m]; vm = C.vol[m]; kzm = C.kz[m]; tzm = C.tz[m]; gzm = C.gz[m];
      }
      V2<CT> x0 = LX(c);
      CT d0 = C.dut[c], v0 = C.vol[c];
      for (int k = k0; k < k1; ++k, c += sz) {
        V2<CT> xp{0, 0};
        CT dp = 0, vp = 0;
        if (k + 1 < nz) { xp = LX(c + sz); dp = C.dut[c + sz]; vp = C.vol[c + sz]; }
        const CT kz0 = C.kz[c], tz0 = C.tz[c], gz0 = C.gz[c];
        CT s1 = fa * v0 * x0.a, s2 = v0 * x0.b;
        if constexpr (TR == 0) {
          const CT q0 = d0 * x0.a;
          auto hf = [&](CT kk, CT th, long n) {
            CT q = C.dut[n] * LX(n).a;
            CT t = kk * (q0 - q);
            s1 += th * t; s2 += t;
          };
          if (i > 0) hf(C.kx[c - 1], C.tx[c - 1], c - 1);
          if (i + 1 < nx) hf(C.kx[c], C.tx[c], c + 1);
          if (j > 0) hf(C.ky[c - sy], C.ty[c - sy], c - sy);
          if (j + 1 < ny) hf(C.ky[c], C.ty[c], c + sy);
          if (k > 0) {
            CT t = kzm * (q0 - dm * xm.a);
            CT B = gzm * (vm * xm.b + v0 * x0.b);
            s1 += tzm * (t + B); s2 += t + B;
          }
          if (k + 1 < nz) {
            CT t = kz0 * (q0 - dp * xp.a);
            CT B = gz0 * (v0 * x0.b + vp * xp.b);
            s1 += tz0 * (t - B); s2 += t - B;
          }
        } else {
          CT h = 0;
          auto hf = [&](CT kk, CT th, long n) {
            V2<CT> xn = LX(n);
            h += kk * ((th * x0.a + x0.b) - (th * xn.a + xn.b));
          };
          if (i > 0) hf(C.kx[c - 1], C.tx[c - 1], c - 1);
          if (i + 1 < nx) hf(C.kx[c], C.tx[c], c + 1);
          if (j > 0) hf(C.ky[c - sy], C.ty[c - sy], c - sy);
          if (j + 1 < ny) hf(C.ky[c], C.ty[c], c + sy);
          if (k > 0) {
            CT zc = tzm * x0.a + x0.b, zn = tzm * xm.a + xm.b;
            h += kzm * (zc - zn);
            s2 += gzm * v0 * (zc - zn);
          }
          if (k + 1 < nz) {
            CT zc = tz0 * x0.a + x0.b, zn = tz0 * xp.a + xp.b;
            h += kz0 * (zc - zn);
            s2 += gz0 * v0 * (zn - zc);
          }
          s1 += d0 * h;
        }
        double y1, y2;
        if (a.b) {
          double2 bb = *reinterpret_cast<const double2*>(a.b + 2 * c);
          y1 = bb.x + (double)s1; y2 = bb.y + (double)s2;
        } else {
          y1 = -(double)s1; y2 = -(double)s2;
        }
        if (a.xw) st2(a.xw, c, (double)x0.a, (double)x0.b);
        st2(a.y, c, y1, y2);
        if (a.y2) st2(a.y2, c, y1, y2);
        if (a.dots & DOT_YY) acc[0] += y1 * y1 + y2 * y2;
        if (a.dots & DOT_YU) {
          V2<VU> uu = ld2(a.u, c);
          acc[1] += y1 * (double)uu.a + y2 * (double)uu.b;
        }
        xm = x0; dm = d0; vm = v0; kzm = kz0; tzm = tz0; gzm = gz0;
        x0 = xp; d0 = dp; v0 = vp;
      }
    }
  }
  if (a.dots) block_reduce_store<2>(acc, a.partial, blk, tid, 256);
}

template <int NP, int TR, typename CT, typename VX, typename VO, typename VU>
static void apply_t(hipStream_t s, const Geom& g, const Coef<CT>& c, const ApplyReq& r) {
  KApply<CT, VX, VO, VU> a;
  a.c = c; a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.fa = (CT)g.fa;
  a.x = (const VX*)r.x; a.dx = r.dx; a.xw = r.xw; a.b = r.b; a.y = (VO*)r.y;
  a.y2 = (VO*)r.y2; a.u = (const VU*)r.u; a.dots = r.dots; a.partial = r.partial;
  a.done = r.done;
  int gx = (g.nx + 63) / 64, gy = (g.ny + 3) / 4;
  int cols = gx * gy;
  int nzc = (2048 + cols - 1) / cols;
  if (nzc < 1) nzc = 1;
  if (nzc > g.nz) nzc = g.nz;
  a.kchunk = (g.nz + nzc - 1) / nzc;
  nzc = (g.nz + a.kchunk - 1) / a.kchunk;
  hipLaunchKernelGGL((k_apply<NP, TR, CT, VX, VO, VU>), dim3(gx, gy, nzc), dim3(64, 4), 0, s, a);
}

template <int NP, int TR>
static void apply_np(hipStream_t s, const Geom& g, const Coef<double>& c64,
                     const Coef<float>& c32, const ApplyReq& r) {
  if (r.prec == 0) {
    if (r.vo == 0) apply_t<NP, TR, double, double, double, double>(s, g, c64, r);
    else apply_t<NP, TR, double, double, float, double>(s, g, c64, r);
  } else {
    apply_t<NP, TR, float, float, float, float>(s, g, c32, r);
  }
}

void launch_apply(hipStream_t s, const Geom& g, const Coef<double>& c64, const Coef<float>& c32,
                  const ApplyReq& r) {
  if (g.np1 == 1) {
    if (r.tr) apply_np<1, 1>(s, g, c64, c32, r); else apply_np<1, 0>(s, g, c64, c32, r);
  } else {
    if (r.tr) apply_np<2, 1>(s, g, c64, c32, r); else apply_np<2, 0>(s, g, c64, c32, r);
  }
}

// ===================================================================== line solve
template <typename CT, typename VT>
struct KLine {
  Coef<CT> c;
  int nx, ny, nz, use_lds;
  const VT* src;
  VT* p;
  const VT* v;
  VT* out;
  const double* scale;
  const KState* st;
  double* partial;
  const int* done;
};

int line_blocks(const Geom& g) { return ((g.nx + 63) / 64) * g.ny; }

// Dynamic LDS a line-solve workgroup may take: the device's per-workgroup limit
// (capped at 64 KiB) minus 1 KiB kept for the static reduction scratch.
static size_t lds_budget() {
  static int cached = -1;
  if (cached < 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      v = 0;
    if (v > 64 * 1024) v = 64 * 1024;
    cached = v > 1024 ? v - 1024 : 0;
  }
  return (size_t)cached;
}

template <int NP, int TR, int MODE, typename CT, typename VT>
__global__ __launch_bounds__(64) void k_line(KLine<CT, VT> a) {
  if (a.done && *a.done) return;
  extern __shared__ unsigned char lds_raw[];
  CT* stage = reinterpret_cast<CT*>(lds_raw);
  const int nx = a.nx, nz = a.nz;
  const long sz = (long)nx * a.ny;
  const int tid = threadIdx.x;
  const int i = blockIdx.x * 64 + tid, j = blockIdx.y;
  const int blk = blockIdx.x + gridDim.x * blockIdx.y;
  double acc[1] = {0.0};
  const Coef<CT>& C = a.c;
  if (i < nx) {
    const long c0 = i + (long)nx * j;
    CT alpha = 0, beta = 0, omega = 0, sc = 1;
    bool first = false;
    if constexpr (MODE == LINE_BICG_P) {
      first = a.st->iter == 1;
      beta = (CT)a.st->beta; omega = (CT)a.st->omega;
    }
    if constexpr (MODE == LINE_BICG_S) alpha = (CT)a.st->alpha;
    if constexpr (MODE == LINE_PLAIN) { if (a.scale) sc = (CT)(*a.scale); }
    if constexpr (NP == 1) {
      auto form = [&](long c) -> CT {
        if constexpr (MODE == LINE_PLAIN) return (CT)a.src[c] * sc;
        else if constexpr (MODE == LINE_BICG_P) {
          CT r = (CT)a.src[c];
          CT pv = first ? r : r + beta * ((CT)a.p[c] - omega * (CT)a.v[c]);
          a.p[c] = (VT)pv;
          return pv;
        } else {
          CT sv = (CT)a.p[c] - alpha * (CT)a.v[c];
          a.p[c] = (VT)sv;
          acc[0] += (double)sv * (double)sv;
          return sv;
        }
      };
      auto put = [&](int k, long c, CT v) {
        if (a.use_lds) stage[k * 64 + tid] = v; else a.out[c] = (VT)v;
      };
      auto get = [&](int k, long c) -> CT {
        return a.use_lds ? stage[k * 64 + tid] : (CT)a.out[c];
      };
      CT yprev = 0;
      long c = c0;
      for (int k = 0; k < nz; ++k, c += sz) {
        CT in = form(c);
        CT y;
        if constexpr (TR == 0) y = (in - C.pl[c] * yprev) * C.pt[c];
        else y = (k > 0) ? in - C.pw[c - sz] * yprev : in;
        put(k, c, y);
        yprev = y;
      }
      CT znext = 0;
      c = c0 + sz * (nz - 1);
      for (int k = nz - 1; k >= 0; --k, c -= sz) {
        CT y = get(k, c), z;
        if constexpr (TR == 0) z = y - C.pw[c] * znext;
        else z = (k + 1 < nz) ? (y - C.pl[c + sz] * znext) * C.pt[c] : y * C.pt[c];
        a.out[c] = (VT)(-z);
        znext = z;
      }
    } else {
      auto form = [&](long c) -> V2<CT> {
        if constexpr (MODE == LINE_PLAIN) {
          V2<VT> s = ld2(a.src, c);
          return {(CT)s.a * sc, (CT)s.b * sc};
        } else if constexpr (MODE == LINE_BICG_P) {
          V2<VT> r = ld2(a.src, c);
          V2<CT> pv{(CT)r.a, (CT)r.b};
          if (!first) {
            V2<VT> pp = ld2((const VT*)a.p, c), vv = ld2(a.v, c);
            pv.a = pv.a + beta * ((CT)pp.a - omega * (CT)vv.a);
            pv.b = pv.b + beta * ((CT)pp.b - omega * (CT)vv.b);
          }
          st2(a.p, c, pv.a, pv.b);
          return pv;
        } else {
          V2<VT> r = ld2((const VT*)a.p, c), vv = ld2(a.v, c);
          V2<CT> sv{(CT)r.a - alpha * (CT)vv.a, (CT)r.b - alpha * (CT)vv.b};
          st2(a.p, c, sv.a, sv.b);
          acc[0] += (double)sv.a * (double)sv.a + (double)sv.b * (double)sv.b;
          return sv;
        }
      };
      auto put = [&](int k, long c, V2<CT> v) {
        if (a.use_lds) { stage[(2 * k) * 64 + tid] = v.a; stage[(2 * k + 1) * 64 + tid] = v.b; }
        else st2(a.out, c, v.a, v.b);
      };
      auto get = [&](int k, long c) -> V2<CT> {
        if (a.use_lds) return {stage[(2 * k) * 64 + tid], stage[(2 * k + 1) * 64 + tid]};
        V2<VT> o = ld2((const VT*)a.out, c);
        return {(CT)o.a, (CT)o.b};
      };
      auto M = [&](const CT* m, long c, V2<CT> v) -> V2<CT> {  // m * v
        const CT* q = m + 4 * c;
        return {q[0] * v.a + q[2] * v.b, q[1] * v.a + q[3] * v.b};
      };
      auto MT = [&](const CT* m, long c, V2<CT> v) -> V2<CT> {  // m^T * v
        const CT* q = m + 4 * c;
        return {q[0] * v.a + q[1] * v.b, q[2] * v.a + q[3] * v.b};
      };
      V2<CT> yprev{0, 0};
      long c = c0;
      for (int k = 0; k < nz; ++k, c += sz) {
        V2<CT> in = form(c), y;
        if constexpr (TR == 0) {
          V2<CT> ly = M(C.pl, c, yprev);
          y = M(C.pt, c, V2<CT>{in.a - ly.a, in.b - ly.b});
        } else {
          if (k > 0) {
            V2<CT> wy = MT(C.pw, c - sz, yprev);
            y = {in.a - wy.a, in.b - wy.b};
          } else {
            y = in;
          }
        }
        put(k, c, y);
        yprev = y;
      }
      V2<CT> znext{0, 0};
      c = c0 + sz * (nz - 1);
      for (int k = nz - 1; k >= 0; --k, c -= sz) {
        V2<CT> y = get(k, c), z;
        if constexpr (TR == 0) {
          V2<CT> wz = M(C.pw, c, znext);
          z = {y.a - wz.a, y.b - wz.b};
        } else {
          V2<CT> t = y;
          if (k + 1 < nz) {
            V2<CT> lz = MT(C.pl, c + sz, znext);
            t = {y.a - lz.a, y.b - lz.b};
          }
          z = MT(C.pt, c, t);
        }
        st2(a.out, c, -z.a, -z.b);
        znext = z;
      }
    }
  }
  if constexpr (MODE == LINE_BICG_S) block_reduce_store<1>(acc, a.partial, blk, tid, 64);
}

template <int NP, int TR, typename CT, typename VT>
static void line_t(hipStream_t s, const Geom& g, const Coef<CT>& c, const LineReq& r) {
  KLine<CT, VT> a;
  a.c = c; a.nx = g.nx; a.ny = g.ny; a.nz = g.nz;
  a.src = (const VT*)r.src; a.p = (VT*)r.p; a.v = (const VT*)r.v; a.out = (VT*)r.out;
  a.scale = r.scale; a.st = r.st; a.partial = r.partial; a.done = r.done;
  size_t lds = (size_t)g.nz * 64 * NP * sizeof(CT);
  a.use_lds = (r.use_lds && lds <= lds_budget()) ? 1 : 0;
  if (!a.use_lds) lds = 0;
  dim3 grid((g.nx + 63) / 64, g.ny);
  switch (r.mode) {
    case LINE_PLAIN:
      hipLaunchKernelGGL((k_line<NP, TR, LINE_PLAIN, CT, VT>), grid, dim3(64), lds, s, a); break;
    case LINE_BICG_P:
      hipLaunchKernelGGL((k_line<NP, TR, LINE_BICG_P, CT, VT>), grid, dim3(64), lds, s, a); break;
    default:
      hipLaunchKernelGGL((k_line<NP, TR, LINE_BICG_S, CT, VT>), grid, dim3(64), lds, s, a); break;
  }
}

template <int NP, int TR>
static void line_np(hipStream_t s, const Geom& g, const Coef<double>& c64,
                    const Coef<float>& c32, const LineReq& r) {
  if (r.prec == 0) line_t<NP, TR, double, double>(s, g, c64, r);
  else line_t<NP, TR, float, float>(s, g, c32, r);
}

void launch_line(hipStream_t s, const Geom& g, const Coef<double>& c64, const Coef<float>& c32,
                 const LineReq& r) {
  if (g.np1 == 1) {
    if (r.tr) line_np<1, 1>(s, g, c64, c32, r); else line_np<1, 0>(s, g, c64, c32, r);
  } else {
    if (r.tr) line_np<2, 1>(s, g, c64, c32, r); else line_np<2, 0>(s, g, c64, c32, r);
  }
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

}  // namespace asam
