// noalloc.hip -- the allocation-free dot product: the partial sums of asam_dev_dot's
// reduction live in a file-scope device array instead of a context's buffers, and the
// result goes to caller-owned device memory, so a call allocates nothing and does not
// synchronise (usable in loops and code that may not allocate device memory).
#include <mutex>

#include "asam_sound.h"
#include "internal.h"
#include "kcommon.h"

namespace asam {

// Partial sums of one reduction: blas_blocks() never launches more than this many
// workgroups, so the array holds every partial of any n (8 KiB per device).
constexpr int kNoAllocSlots = 1024;
static __device__ __align__(256) double scratch[kNoAllocSlots];

constexpr int kMaxDevices = 64;
static double* g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Device address of this device's copy of `scratch` (looked up once per device, with the
// device current).  Host code uses only this address, never the array's own name.
static int scratch_for_current_device(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return ASAM_E_HIP;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return ASAM_E_HIP;
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

// Sum of nblk partials in k_finalize's fixed order (strided per-thread sums, wavefront
// butterfly, the 16 wave totals in order), so the result has asam_dev_dot's bits.
__global__ __launch_bounds__(1024) void k_sum_partials(int nblk, const double* partial,
                                                       double* out) {
  __shared__ double sh[16];
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int b = tid; b < nblk; b += 1024) s += partial[b];
  s = wave_sum(s);
  if ((tid & 63) == 0) sh[tid >> 6] = s;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int w = 0; w < 16; ++w) t += sh[w];
    out[0] = t;
  }
}

}  // namespace asam

using namespace asam;

extern "C" int asam_dev_dot_noalloc(void* stream, size_t n, const double* x, const double* y,
                                    double* dot) {
  if (!x || !y || !dot) return ASAM_E_INVALID;
  if (n > (size_t)ASAM_DOT_NOALLOC_MAX_N) return ASAM_E_INVALID;
  double* part = nullptr;
  int rc = scratch_for_current_device(&part);
  if (rc) return rc;
  const int nblk = blas_blocks((long)n);
  if (nblk > kNoAllocSlots) return ASAM_E_INVALID;
  hipStream_t s = static_cast<hipStream_t>(stream);
  launch_dot(s, (long)n, x, y, part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, nblk, (const double*)part, dot);
  if (hipGetLastError() != hipSuccess) return ASAM_E_HIP;
  return 0;
}
