// Vector kernels of GMRES-based iterative refinement (GMRES-IR) on gfx950:
// block dot products, basis combinations and column scaling over global
// vectors with many right-hand sides (ndof x n_rhs, column major; a basis
// of k such blocks at a fixed stride).  Every reduction runs in a fixed
// order (per-workgroup partial sums, then one ordered final pass), so a
// refinement gives the same bits on every run.
//
// These kernels serve the refinement slot of Mesh_::improve_chol /
// improve_lu_piv (UHM/src/mesh/chol.cxx:136-149, UHM/src/mesh/lu_piv.cxx:232-245),
// whose reference body is a stub.
#include <hip/hip_runtime.h>

#include <mutex>

#include "uhmk.h"

namespace {

constexpr int KMAX = 32;   // basis blocks per dot-product launch
constexpr int ROWS = 4096; // rows per workgroup of the dot products

// partial[(blk * KMAX + i) * nrhs + c] = sum over the block's rows of
// V_i(g, c) w(g, c), i < k.  grid (row blocks, nrhs).
__global__ __launch_bounds__(256) void k_block_dots(const double* __restrict__ V, long long vstride, int k,
                                                    const double* __restrict__ w, long long nrow, int nrhs,
                                                    double* __restrict__ partial) {
  __shared__ double red[4][KMAX];
  const int c = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const long long g0 = (long long)blockIdx.x * ROWS;
  const long long g1 = g0 + ROWS < nrow ? g0 + ROWS : nrow;
  const long long co = (long long)c * nrow;
  double acc[KMAX];
#pragma unroll
  for (int i = 0; i < KMAX; ++i) acc[i] = 0.0;
  for (long long g = g0 + tid; g < g1; g += 256) {
    const double x = w[co + g];
#pragma unroll
    for (int i = 0; i < KMAX; ++i)
      if (i < k) acc[i] = fma(V[i * vstride + co + g], x, acc[i]);
  }
#pragma unroll
  for (int i = 0; i < KMAX; ++i) {
    if (i < k) {
      double a = acc[i];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
      if (lane == 0) red[wv][i] = a;
    }
  }
  __syncthreads();
  if (tid < k) {
    const double s = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
    partial[((long long)blockIdx.x * KMAX + tid) * nrhs + c] = s;
  }
}

// out[i * nrhs + c] = sum over row blocks (in order) of the partials.
__global__ __launch_bounds__(256) void k_block_dots_final(const double* __restrict__ partial, int nblk, int k,
                                                          int nrhs, double* __restrict__ out) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= k * nrhs) return;
  const int i = idx / nrhs, c = idx % nrhs;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partial[((long long)b * KMAX + i) * nrhs + c];
  out[idx] = s;
}

// w(g, c) = beta w(g, c) + alpha sum_{i < k} coef[i * nrhs + c] V_i(g, c)
// (beta 1: update; beta 0: w is overwritten).
// grid (row blocks, nrhs): one column per blockIdx.y.
__global__ __launch_bounds__(256) void k_basis_combine(const double* __restrict__ V, long long vstride, int k,
                                                       const double* __restrict__ coef, double alpha, double beta,
                                                       double* __restrict__ w, long long nrow, int nrhs) {
  const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
  if (g >= nrow) return;
  const int c = blockIdx.y;
  const long long idx = (long long)c * nrow + g;
  double s = 0.0;
  for (int i = 0; i < k; ++i) s = fma(coef[i * nrhs + c], V[i * vstride + idx], s);
  w[idx] = beta == 0.0 ? alpha * s : fma(alpha, s, beta * w[idx]);
}

// w(:, c) *= s[c]
__global__ __launch_bounds__(256) void k_scale_columns(double* __restrict__ w, long long nrow, int nrhs,
                                                       const double* __restrict__ s) {
  const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
  if (g >= nrow) return;
  w[(long long)blockIdx.y * nrow + g] *= s[blockIdx.y];
}

inline int ret(hipError_t e) { return e == hipSuccess ? 0 : -(int)e; }

// Partial sums of uhmk_block_dots_noalloc: a fixed device array instead of
// a caller's workspace, for callers that cannot allocate.  Every GPU holds
// its own copy; the host reaches it only through the address that
// hipGetSymbolAddress returns on that device (kept per device below).
constexpr long long NOALLOC_DOUBLES = 1LL << 19;  // 4 MiB
__device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int MAX_DEVICES = 64;
double* g_scratch_ptr[MAX_DEVICES] = {};
std::once_flag g_scratch_once[MAX_DEVICES];
hipError_t g_scratch_err[MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current; nullptr (and *err set) on failure.
double* scratch_ptr(hipError_t* err) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    *err = e;
    return nullptr;
  }
  if (dev < 0 || dev >= MAX_DEVICES) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::call_once(g_scratch_once[dev], [dev] {
    void* p = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (g_scratch_err[dev] == hipSuccess) g_scratch_ptr[dev] = static_cast<double*>(p);
  });
  *err = g_scratch_err[dev];
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

long long uhmk_block_dots_scratch(long long nrow, int nrhs) {
  return ((nrow + ROWS - 1) / ROWS) * (long long)KMAX * nrhs;
}

int uhmk_block_dots(const double* V, long long vstride, int k, const double* w, long long nrow, int nrhs,
                    double* partial, double* out, hipStream_t stream) {
  if (k <= 0 || nrow <= 0 || nrhs <= 0) return 0;
  if (k > KMAX || nrhs > 65535) return -1012;
  const long long nblk = (nrow + ROWS - 1) / ROWS;
  if (nblk > 0x7fffffffLL) return -1012;
  hipLaunchKernelGGL(k_block_dots, dim3((unsigned)nblk, (unsigned)nrhs), dim3(256), 0, stream, V, vstride, k, w,
                     nrow, nrhs, partial);
  hipLaunchKernelGGL(k_block_dots_final, dim3((unsigned)((k * nrhs + 255) / 256)), dim3(256), 0, stream, partial,
                     (int)nblk, k, nrhs, out);
  return ret(hipGetLastError());
}

long long uhmk_block_dots_noalloc_max(void) { return NOALLOC_DOUBLES; }

int uhmk_block_dots_noalloc(const double* V, long long vstride, int k, const double* w, long long nrow, int nrhs,
                            double* out, hipStream_t stream) {
  if (k <= 0 || nrow <= 0 || nrhs <= 0) return 0;
  if (k > KMAX || nrhs > 65535) return -1012;
  if (uhmk_block_dots_scratch(nrow, nrhs) > NOALLOC_DOUBLES) return -1013;
  hipError_t err = hipSuccess;
  double* partial = scratch_ptr(&err);
  if (!partial) return ret(err != hipSuccess ? err : hipErrorInvalidSymbol);
  // every partial slot that k_block_dots_final reads is written first by
  // k_block_dots in the same stream order, so nothing needs clearing
  return uhmk_block_dots(V, vstride, k, w, nrow, nrhs, partial, out, stream);
}

int uhmk_basis_combine(const double* V, long long vstride, int k, const double* coef, double alpha, double beta,
                       double* w, long long nrow, int nrhs, hipStream_t stream) {
  if (nrow <= 0 || nrhs <= 0) return 0;
  if (nrhs > 65535) return -1012;
  hipLaunchKernelGGL(k_basis_combine, dim3((unsigned)((nrow + 255) / 256), (unsigned)nrhs), dim3(256), 0, stream, V,
                     vstride, k, coef, alpha, beta, w, nrow, nrhs);
  return ret(hipGetLastError());
}

int uhmk_scale_columns(double* w, long long nrow, int nrhs, const double* s, hipStream_t stream) {
  if (nrow <= 0 || nrhs <= 0) return 0;
  if (nrhs > 65535) return -1012;
  hipLaunchKernelGGL(k_scale_columns, dim3((unsigned)((nrow + 255) / 256), (unsigned)nrhs), dim3(256), 0, stream, w,
                     nrow, nrhs, s);
  return ret(hipGetLastError());
}

}  // extern "C"
