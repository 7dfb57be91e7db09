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
// Cholesky of the kb x kb diagonal panel block, one workgroup of four
// wavefronts per front (uhmd::CholColStep, fully unrolled), followed by the
// inverse W = L11^-1.  W's strict lower part is stored transposed in the
// strict upper triangle of the diagonal block, which the Cholesky path never
// reads otherwise (only the lower triangle of a front is used); the panel
// TRSM and the solves read it from there (its diagonal is 1 / L(k, k)).
// ------------------------------------------------------------------------
constexpr int LDW = TILE + 1;
typedef double d4 __attribute__((ext_vector_type(4)));

// One wavefront: C += X Y for 16 x 16 blocks X, Y in LDS (row strides ldx,
// ldy), K = 16 as four v_mfma_f64_16x16x4f64.  Operand lanes: X(i, k) in lane
// i + 16 k, Y(k, j) in lane j + 16 k; result register r of lane l holds
// C(4 r + l / 16, l % 16).
__device__ __forceinline__ d4 mma16_ld(const double* X, int ldx, const double* Y, int ldy, d4 c, int lane) {
#pragma unroll
  for (int k0 = 0; k0 < 16; k0 += 4) {
    const int k = k0 + (lane >> 4), q = lane & 15;
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(X[q * ldx + k], Y[k * ldy + q], c, 0, 0, 0);
  }
  return c;
}
__device__ __forceinline__ d4 mma16(const double* X, const double* Y, int ld, d4 c, int lane) {
  return mma16_ld(X, ld, Y, ld, c, lane);
}

template <typename T>
__global__ __launch_bounds__(256) void k_chol_diag(const uhmk_front* __restrict__ F,
                                                   const int* __restrict__ flist, int c0, int nb,
                                                   T* __restrict__ arena, int* __restrict__ info) {
  __shared__ double sc[2][64];
  __shared__ double sL[TILE * LDW];
  __shared__ double sW[TILE * LDW];
  __shared__ double sS[4][16 * 17];
  __shared__ int s_first[4];
  const int f = flist[blockIdx.x];
  const uhmk_front fr = F[f];
  const int n = fr.n, fs = fr.fs;
  const int kb = min(nb, fs - c0);
  T* __restrict__ A = arena + fr.off + c0 + (long long)c0 * n;
  const int i = threadIdx.x & 63, w = threadIdx.x >> 6;
  double a[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int j = w + 4 * u;
    a[u] = uhmd::ldz(A, i + (long long)j * n, i < kb && j < kb && j <= i);
  }
  int first = 0;
  uhmd::CholColStep<0>::run(a, i, w, kb, first, sc);
  if (i == 0) s_first[w] = first;
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int j = w + 4 * u;
    sL[i * LDW + j] = a[u];
    if (i < kb && j < kb && j <= i) A[i + (long long)j * n] = (T)a[u];
  }
  __syncthreads();
  // W = L^-1 by 16 x 16 blocks: the four diagonal blocks are inverted in
  // parallel (one wavefront each, lane c < 16 forms column c), then the
  // blocks below the diagonal follow diagonal by diagonal,
  // W(I,J) = -W(I,I) sum_{K=J}^{I-1} L(I,K) W(K,J), as 16 x 16 x 16 FP64
  // MFMA products through LDS.
  {
    const int b0 = 16 * w;
    if (i < 16) {
      double wv[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        double acc = (k == i) ? 1.0 : 0.0;
#pragma unroll
        for (int m = 0; m < k; ++m) acc = fma(-sL[(b0 + k) * LDW + b0 + m], wv[m], acc);
        const double d = sL[(b0 + k) * LDW + b0 + k];
        wv[k] = (b0 + k < kb) ? acc / d : 0.0;
      }
#pragma unroll
      for (int k = 0; k < 16; ++k) sW[(b0 + k) * LDW + b0 + i] = wv[k];
    }
  }
  __syncthreads();
  for (int d = 1; d < 4; ++d) {
    // wavefront w forms block (I, J) = (w + d, w) when it exists
    const int I = w + d, J = w;
    d4 S = d4{0.0, 0.0, 0.0, 0.0};
    if (I < 4) {
      // S = sum_K L(I,K) W(K,J)
      for (int K = J; K < I; ++K) S = mma16(sL + (16 * I) * LDW + 16 * K, sW + (16 * K) * LDW + 16 * J, LDW, S, i);
#pragma unroll
      for (int r = 0; r < 4; ++r) sS[w][(4 * r + (i >> 4)) * 17 + (i & 15)] = S[r];
    }
    __syncthreads();
    if (I < 4) {
      d4 Wb = d4{0.0, 0.0, 0.0, 0.0};
      Wb = mma16_ld(sW + (16 * I) * LDW + 16 * I, LDW, sS[w], 17, Wb, i);
#pragma unroll
      for (int r = 0; r < 4; ++r) sW[(16 * I + 4 * r + (i >> 4)) * LDW + 16 * J + (i & 15)] = -Wb[r];
    }
    __syncthreads();
  }
  double wc[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) wc[u] = sW[i * LDW + w + 4 * u];
  // strict lower part of W, transposed into the strict upper triangle
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int j = w + 4 * u;
    if (i < kb && j < i) A[j + (long long)i * n] = (T)wc[u];
  }
  __syncthreads();
  if (threadIdx.x == 0 && info[f] == 0) {
    int k1 = 0;
    for (int q = 0; q < 4; ++q)
      if (s_first[q] && (k1 == 0 || s_first[q] < k1)) k1 = s_first[q];
    if (k1) info[f] = c0 + k1;
  }
}

// Cholesky panel rows [b, b+64) below the diagonal block: X <- X L11^-T =
// X W^T with W = L11^-1 from the diagonal block (strict upper triangle:
// W transposed; diagonal: 1 / L(k,k)).  One 64 x 64 x kb FP64 MFMA product
// per workgroup; each wavefront owns a 32 x 32 quarter of the result.
constexpr int LDT = TILE + 4;

template <typename T>
__global__ __launch_bounds__(256) void k_chol_trsm(const uhmk_front* __restrict__ F,
                                                   const uhmk_item* __restrict__ items, int c0, int nb,
                                                   T* __restrict__ arena) {
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int kb = min(nb, fs - c0);
  T* __restrict__ A = arena + fr.off;
  __shared__ double sWt[TILE * LDT];  // sWt[k][j] = W(j, k)
  __shared__ double sXt[TILE * LDT];  // sXt[k][i] = X(i, k)
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const T* __restrict__ D = A + c0 + (long long)c0 * n;
  double vw[16], vx[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int idx = tid + 256 * u, r = idx & 63, c = idx >> 6;
    vw[u] = uhmd::ldz(D, r + (long long)c * n, r < kb && c < kb && r <= c);
    const int row = it.b + r;
    vx[u] = uhmd::ldz(A, row + (long long)(c0 + c) * n, row < n && c < kb);
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int idx = tid + 256 * u, r = idx & 63, c = idx >> 6;
    // block entry (r, c), r <= c: W(c, r) (r < c) or 1 / L(c, c) (r == c)
    sWt[r * LDT + c] = (r == c && r < kb) ? 1.0 / vw[u] : vw[u];
    sXt[c * LDT + r] = vx[u];
  }
  __syncthreads();
  const int iw = (w & 1) * 32, jw = (w >> 1) * 32, cl = lane & 15;
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k0 = 0; k0 < TILE; k0 += 4) {
    const int kr = k0 + (lane >> 4);
    double wo[2], xo[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) wo[b] = sWt[kr * LDT + jw + 16 * b + cl];
#pragma unroll
    for (int a = 0; a < 2; ++a) xo[a] = sXt[kr * LDT + iw + 16 * a + cl];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(wo[b], xo[a], acc[a][b], 0, 0, 0);
  }
  // accumulator register r of lane l: D(j, i) with j = 4 r + l / 16, i = l % 16
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int row = it.b + iw + 16 * a + cl;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = jw + 16 * b + 4 * r + (lane >> 4);
        if (row < n && j < kb) A[row + (long long)(c0 + j) * n] = (T)acc[a][b][r];
      }
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
__global__ __launch_bounds__(256, 2) void k_gemm(const uhmk_front* __restrict__ F,
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
      ra[u] = uhmd::ldz_as<TM>(M, (i0 + li) + (long long)k * n, i0 + li < i1 && k < kend);
    }
    if (TRANSB) {
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int k = k0 + lk + KR * u;
        rb[u] = uhmd::ldz_as<TM>(M, (j0 + li) + (long long)k * n, j0 + li < j1 && k < kend);
      }
    } else {
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int j = bj + 16 * u;
        rb[u] = uhmd::ldz_as<TM>(M, (k0 + bk) + (long long)(j0 + j) * n, j0 + j < j1 && k0 + bk < kend);
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
    hipLaunchKernelGGL(k_chol_diag<double>, dim3(n_fronts), dim3(256), 0, stream, fronts, flist, c0, nb,
                       static_cast<double*>(arena), info);
  else if (prec == 1)
    hipLaunchKernelGGL(k_chol_diag<float>, dim3(n_fronts), dim3(256), 0, stream, fronts, flist, c0, nb,
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
