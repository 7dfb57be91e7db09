// Dense front kernels for gfx950 (FP64): Cholesky diagonal block, panel
// triangular solves and the MFMA Schur / trailing update.
//
// Reference operations restated (semantics, not code):
//   FLA_Chol on the diagonal tile      LINAL/src/dense/chol.cxx:67
//   trsm(R, L, T) on the panel rows    LINAL/src/dense/chol.cxx:80-84,
//                                      UHM/src/matrix/uhm/fla/chol/decompose.cxx:93-94
//   gemm / syrk updates                LINAL/src/dense/gemm.cxx:48-137, syrk.cxx:43-123,
//                                      lu_piv/decompose.cxx:148-149, chol/decompose.cxx:95-96
#include <hip/hip_runtime.h>

#include "dense_device.hpp"
#include "uhmk.h"

namespace {

constexpr int TILE = 64;
constexpr int LDP = TILE + 1;

// ------------------------------------------------------------------------
// Cholesky of the kb x kb diagonal panel block, one wavefront per front:
// lane i holds row i in registers (uhmd::CholStep, fully unrolled).
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64) void k_chol_diag(const uhmk_front* __restrict__ F,
                                                  const int* __restrict__ flist, int c0, int nb,
                                                  T* __restrict__ arena, int* __restrict__ info) {
  const int f = flist[blockIdx.x];
  const uhmk_front fr = F[f];
  const int n = fr.n, fs = fr.fs;
  const int kb = min(nb, fs - c0);
  T* __restrict__ A = arena + fr.off;
  const int i = threadIdx.x;
  double a[TILE];
#pragma unroll
  for (int j = 0; j < TILE; ++j)
    a[j] = (i < kb && j < kb && j <= i) ? (double)A[(c0 + i) + (long long)(c0 + j) * n] : 0.0;
  int first = 0;
  uhmd::CholStep<0>::run(a, i, kb, first);
  if (i == 0 && first && info[f] == 0) info[f] = c0 + first;
#pragma unroll
  for (int j = 0; j < TILE; ++j)
    if (i < kb && j < kb && j <= i) A[(c0 + i) + (long long)(c0 + j) * n] = (T)a[j];
}

// Cholesky panel rows [i0, i0+64) below the diagonal block: X <- X L11^-T.
// 16 rows per wavefront, 4 lanes per row (tri_solve64).
template <typename T>
__global__ __launch_bounds__(256) void k_chol_trsm(const uhmk_front* __restrict__ F,
                                                   const uhmk_item* __restrict__ items, int c0, int nb,
                                                   T* __restrict__ arena) {
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int kb = min(nb, fs - c0);
  T* __restrict__ A = arena + fr.off;
  __shared__ double s_T[TILE * LDP];
  __shared__ double s_rd[TILE];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  uhmd::load_tri64<true>(s_T, LDP, s_rd, A + c0 + (long long)c0 * n, n, kb, tid, 256);
  __syncthreads();
  const int row = it.b + 16 * w + (lane >> 2), q = lane & 3;
  const bool live = row < n;
  double x[16];
#pragma unroll
  for (int cr = 0; cr < 16; ++cr) {
    const int c = 16 * q + cr;
    x[cr] = (live && c < kb) ? (double)A[row + (long long)(c0 + c) * n] : 0.0;
  }
  uhmd::tri_solve64<false, false>(x, s_T, LDP, s_rd, kb, lane);
  if (live)
#pragma unroll
    for (int cr = 0; cr < 16; ++cr) {
      const int c = 16 * q + cr;
      if (c < kb) A[row + (long long)(c0 + c) * n] = (T)x[cr];
    }
}

// ------------------------------------------------------------------------
// Schur update C -= A * op(B) on T x T tiles (T = 64 or 128) on MFMA.
// A = F[i, k], B = F[k, j] (NN) or F[j, k] (NT), C = F[i, j]; K = [c0,
// min(c0+nb, fs)) for a panel update or [0, fs) for the final Schur
// complement (schur = 1).  Tiles never straddle fs, so the extent is
// min(i0 + T, fs) or min(i0 + T, n).  Four waves each own a (T/2) x (T/2)
// quarter; K advances in chunks of 16 through LDS while the next chunk is
// prefetched into registers.  TS is the storage type of the fronts, TM the
// MFMA operand type: double -> v_mfma_f64_16x16x4f64, float ->
// v_mfma_f32_16x16x4f32 (FP32 products and accumulation, operands rounded
// to FP32 on their way to LDS).
// Operand layout of both 16x16x4 instructions: A[i][k] in lane i + 16 k,
// B[k][j] in lane j + 16 k.  Accumulator register r of lane l holds
// D[4 r + l / 16][l % 16] (f64) or D[4 (l / 16) + r][l % 16] (f32).  The
// product is formed transposed (D = B^T A^T) so that a 16-lane group
// stores 16 consecutive rows of C.
// ------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int KC = 16;

template <typename TM>
struct Mfma;
template <>
struct Mfma<double> {
  typedef d4 acc_t;
  __device__ __forceinline__ static d4 run(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ __forceinline__ static int col(int r, int lane) { return 4 * r + (lane >> 4); }
};
template <>
struct Mfma<float> {
  typedef f4 acc_t;
  __device__ __forceinline__ static f4 run(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ __forceinline__ static int col(int r, int lane) { return 4 * (lane >> 4) + r; }
};

template <bool TRANSB, int T, typename TS, typename TM>
__global__ __launch_bounds__(256) void k_gemm(const uhmk_front* __restrict__ F,
                                              const uhmk_item* __restrict__ items, int c0, int nb, int schur,
                                              TS* __restrict__ arena) {
  typedef typename Mfma<TM>::acc_t acc_t;
  constexpr int LD = T + 16;          // row stride of the LDS panels (conflict-free half-wave reads)
  constexpr int NB = T / 32;          // 16x16 blocks per wave and dimension
  constexpr int KR = 256 / T;         // k rows per load pass (one thread per i)
  constexpr int NL = KC / KR;         // loads per thread per operand
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int kbeg = schur ? 0 : c0;
  const int kend = schur ? fs : min(c0 + nb, fs);
  const int i0 = it.a, j0 = it.b;
  const int i1 = min(i0 + T, i0 < fs ? fs : n);
  const int j1 = min(j0 + T, j0 < fs ? fs : n);
  TS* __restrict__ M = arena + fr.off;
  __shared__ TM sA[KC * LD];
  __shared__ TM sB[KC * LD];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int iw = (w & 1) * (T / 2), jw = (w >> 1) * (T / 2);

  acc_t acc[NB][NB];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

  TM ra[NL], rb[NL];
  // A (and NT B): thread t loads row i = t % T of k rows t / T + KR u.
  const int li = tid % T, lk = tid / T;
  // NN B: thread t loads k = t % 16 of columns t / 16 + 16 u.
  const int bk = tid & 15, bj = tid >> 4;
  auto load = [&](int k0) {
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int k = k0 + lk + KR * u;
      ra[u] = (i0 + li < i1 && k < kend) ? (TM)M[(i0 + li) + (long long)k * n] : (TM)0;
    }
    if (TRANSB) {
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int k = k0 + lk + KR * u;
        rb[u] = (j0 + li < j1 && k < kend) ? (TM)M[(j0 + li) + (long long)k * n] : (TM)0;
      }
    } else {
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int j = bj + 16 * u;
        rb[u] = (j0 + j < j1 && k0 + bk < kend) ? (TM)M[(k0 + bk) + (long long)(j0 + j) * n] : (TM)0;
      }
    }
  };
  if (kbeg < kend) load(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += KC) {
#pragma unroll
    for (int u = 0; u < NL; ++u) sA[(lk + KR * u) * LD + li] = ra[u];
    if (TRANSB) {
#pragma unroll
      for (int u = 0; u < NL; ++u) sB[(lk + KR * u) * LD + li] = rb[u];
    } else {
#pragma unroll
      for (int u = 0; u < NL; ++u) sB[bk * LD + bj + 16 * u] = rb[u];
    }
    __syncthreads();
    if (k0 + KC < kend) load(k0 + KC);
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
      const int kr = kk + (lane >> 4), cl = lane & 15;
      TM bo[NB], ao[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) bo[b] = sB[kr * LD + jw + b * 16 + cl];
#pragma unroll
      for (int a = 0; a < NB; ++a) ao[a] = sA[kr * LD + iw + a * 16 + cl];
#pragma unroll
      for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = Mfma<TM>::run(bo[b], ao[a], acc[a][b]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int i = i0 + iw + a * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = j0 + jw + b * 16 + Mfma<TM>::col(r, lane);
        if (i < i1 && j < j1) {
          TS* c = M + i + (long long)j * n;
          *c = (TS)((double)*c - (double)acc[a][b][r]);
        }
      }
    }
}

template <typename TS, typename TM>
int launch_gemm(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, int schur,
                int transb, int tile, TS* arena, hipStream_t stream) {
  dim3 g(n_items), b(256);
  if (tile == 128) {
    if (transb) hipLaunchKernelGGL((k_gemm<true, 128, TS, TM>), g, b, 0, stream, fronts, items, c0, nb, schur, arena);
    else hipLaunchKernelGGL((k_gemm<false, 128, TS, TM>), g, b, 0, stream, fronts, items, c0, nb, schur, arena);
  } else if (tile == 64) {
    if (transb) hipLaunchKernelGGL((k_gemm<true, 64, TS, TM>), g, b, 0, stream, fronts, items, c0, nb, schur, arena);
    else hipLaunchKernelGGL((k_gemm<false, 64, TS, TM>), g, b, 0, stream, fronts, items, c0, nb, schur, arena);
  } else {
    return -1004;
  }
  return 0;
}

inline int ret(hipError_t e) { return e == hipSuccess ? 0 : -(int)e; }

}  // namespace

extern "C" {

int uhmk_chol_diag(const uhmk_front* fronts, const int* flist, int n_fronts, int c0, int nb, void* arena, int prec,
                   int* info, hipStream_t stream) {
  if (n_fronts <= 0) return 0;
  if (nb > TILE || nb < 1) return -1000;
  if (prec == 0)
    hipLaunchKernelGGL(k_chol_diag<double>, dim3(n_fronts), dim3(64), 0, stream, fronts, flist, c0, nb,
                       static_cast<double*>(arena), info);
  else if (prec == 1)
    hipLaunchKernelGGL(k_chol_diag<float>, dim3(n_fronts), dim3(64), 0, stream, fronts, flist, c0, nb,
                       static_cast<float*>(arena), info);
  else
    return -1006;
  return ret(hipGetLastError());
}

int uhmk_chol_trsm(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, void* arena,
                   int prec, hipStream_t stream) {
  if (n_items <= 0) return 0;
  if (nb > TILE || nb < 1) return -1000;
  if (prec == 0)
    hipLaunchKernelGGL(k_chol_trsm<double>, dim3(n_items), dim3(256), 0, stream, fronts, items, c0, nb,
                       static_cast<double*>(arena));
  else if (prec == 1)
    hipLaunchKernelGGL(k_chol_trsm<float>, dim3(n_items), dim3(256), 0, stream, fronts, items, c0, nb,
                       static_cast<float*>(arena));
  else
    return -1006;
  return ret(hipGetLastError());
}

int uhmk_gemm_update(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, int schur,
                     int transb, int tile, void* arena, int prec, int mma, hipStream_t stream) {
  if (n_items <= 0) return 0;
  int rc;
  if (prec == 0 && mma == 0)
    rc = launch_gemm<double, double>(fronts, items, n_items, c0, nb, schur, transb, tile, static_cast<double*>(arena),
                                     stream);
  else if (prec == 0 && mma == 1)
    rc = launch_gemm<double, float>(fronts, items, n_items, c0, nb, schur, transb, tile, static_cast<double*>(arena),
                                    stream);
  else if (prec == 1 && mma == 1)
    rc = launch_gemm<float, float>(fronts, items, n_items, c0, nb, schur, transb, tile, static_cast<float*>(arena),
                                   stream);
  else
    return -1006;
  if (rc) return rc;
  return ret(hipGetLastError());
}

}  // extern "C"
