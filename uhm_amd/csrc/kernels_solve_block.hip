// Front solves with many right-hand sides for gfx950: the level solves of
// kernels_solve.hip recast as FP64 MFMA products over tiles of 64
// right-hand sides, so that every factor entry read from HBM feeds 64
// columns of the solution.
//
// Reference operations restated (semantics, not code):
//   forward   solve_lu_piv_1_hier / solve_chol_1_hier
//             UHM/src/matrix/uhm/fla/lu_piv/solve.cxx:161-187,
//             UHM/src/matrix/uhm/fla/chol/solve.cxx:152-168
//   backward  solve_lu_piv_2_hier / solve_chol_2_hier
//             lu_piv/solve.cxx:211-231, chol/solve.cxx:186-202
// with the dense triangular solves and products of LINAL
// (LINAL/src/dense/trsm.cxx, gemm.cxx:48-137) on n x n_rhs blocks.
//
// Per tree level the forward pass is
//   gather  (kernels_solve.hip)   x_f = b_f + children's xb
//   step k0 = 0, 64, ..           y = T_kk^-1 x_k -> z_k;  x_r -= F(r, k) y
//                                 for the 64-row chunks r of ATL below k
//   gemm                          xb -= ABL z_t
// and the backward pass, root first,
//   branch  (kernels_solve.hip)   xb <- the parent's solution rows
//   gemm                          z_t -= ATR xb (LU) or ABL^T xb (Cholesky)
//   step k0 = .., 64, 0           y = T_kk^-1 z_k -> x_k;  z_c -= F(c, k) y
//                                 (LU) or F(k, c)^T y (Cholesky) for the
//                                 64-row chunks c of ATL above k
// The final solution is in x (front layout, n x n_rhs per front, leading
// dimension n); z holds the intermediate ATL rows.  T_kk is the unit lower
// L (LU forward), U (LU backward), L or L^T (Cholesky); for Cholesky factors
// whose 64 x 64 diagonal blocks carry W = L_kk^-1 (uhmk_chol_diag) the
// block step is the product W x_k or W^T z_k on MFMA, otherwise a
// triangular solve of 16 vectors per wavefront (uhmd::TriStep).  Factor
// entries are FP64 or FP32 (the level's storage); arithmetic is FP64.
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "dense_device.hpp"
#include "uhmk.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int NC = 64;   // right-hand sides per tile
constexpr int LD = 68;   // row stride (doubles) of the 64-wide LDS blocks

// acc(c, i) += sum_{k < K} Bm[k][c] Am[k][i] for the wavefront's 32 x 32
// quarter of a 64 x 64 tile: c blocks 2 (w & 1) + {0, 1}, i blocks
// 2 (w >> 1) + {0, 1}.  Bm (right-hand sides) and Am (matrix, transposed)
// live in LDS with row stride LD.  v_mfma_f64_16x16x4f64 operands: lane
// l = m + 16 k holds Bm[k][16 cb + m] (first) and Am[k][16 ib + m]
// (second); result register r of lane l holds acc(c = 16 cb + 4 r + l / 16,
// i = 16 ib + l % 16).
__device__ __forceinline__ void mma_quarter(d4 (&acc)[2][2], const double* __restrict__ Bm,
                                            const double* __restrict__ Am, int K, int lane, int w) {
  const int cb = 32 * (w & 1), ib = 32 * (w >> 1);
  const int q = lane & 15, kq = lane >> 4;
#pragma unroll 4
  for (int k0 = 0; k0 < K; k0 += 4) {
    const double* rb = Bm + (k0 + kq) * LD + cb + q;
    const double* ra = Am + (k0 + kq) * LD + ib + q;
    const double b0 = rb[0], b1 = rb[16];
    const double a0 = ra[0], a1 = ra[16];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a1, acc[1][1], 0, 0, 0);
  }
}

__device__ __forceinline__ void zero_acc(d4 (&acc)[2][2]) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
}

// C(r + i, c0 + c) -= acc(c, i) (STORE: = acc(c, i)) for rows r + i <
// rend and columns c0 + c < nrhs; C has leading dimension n.
template <bool STORE = false>
__device__ __forceinline__ void sub_acc(double* __restrict__ C, long long n, int r, int rend, int c0, int nrhs,
                                        const d4 (&acc)[2][2], int lane, int w) {
  const int cb = 32 * (w & 1), ib = 32 * (w >> 1);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int c = cb + 16 * a + 4 * g + (lane >> 4), i = ib + 16 * b + (lane & 15);
        if (r + i < rend && c0 + c < nrhs) {
          double* p = C + (r + i) + (long long)(c0 + c) * n;
          if (STORE) *p = acc[a][b][g];
          else *p -= acc[a][b][g];
        }
      }
}

// One 64 x 64 tile of P = op(A)[rows r0.., K] S[K, c0..] accumulated into
// acc, K = [kbeg, kend) in 16-deep stages through double-buffered LDS (the
// next stage's global loads are in flight while the current one is
// multiplied).  op(A)(i, k) = A[i + k lda] (NN) or A[k + i lda] (TA); rows
// at or beyond rend and columns at or beyond nrhs read as zero.  S points
// at column c0 (leading dimension lds).
template <typename T, bool TA>
__device__ __forceinline__ void tile_product(d4 (&acc)[2][2], const T* __restrict__ A, long long lda, int r0,
                                             int rend, int kbeg, int kend, const double* __restrict__ S,
                                             long long lds, int c0, int nrhs, double (*sA)[16 * LD],
                                             double (*sB)[16 * LD]) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // staging coordinates: A NN i = tid & 63, k = tid >> 6 (+ 4 u);
  // A TN and S: k = tid & 15, i or c = tid >> 4 (+ 16 u)
  double pa[4], pb[4];
  auto load = [&](int k) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (!TA) {
        const int i = tid & 63, kk = (tid >> 6) + 4 * u;
        pa[u] = uhmd::ldz(A, (r0 + i) + (long long)(k + kk) * lda, r0 + i < rend && k + kk < kend);
      } else {
        const int kk = tid & 15, i = (tid >> 4) + 16 * u;
        pa[u] = uhmd::ldz(A, (k + kk) + (long long)(r0 + i) * lda, r0 + i < rend && k + kk < kend);
      }
      const int kk = tid & 15, c = (tid >> 4) + 16 * u;
      pb[u] = uhmd::ldz(S, (k + kk) + (long long)c * lds, k + kk < kend && c0 + c < nrhs);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (!TA) {
        const int i = tid & 63, kk = (tid >> 6) + 4 * u;
        sA[buf][kk * LD + i] = pa[u];
      } else {
        const int kk = tid & 15, i = (tid >> 4) + 16 * u;
        sA[buf][kk * LD + i] = pa[u];
      }
      const int kk = tid & 15, c = (tid >> 4) + 16 * u;
      sB[buf][kk * LD + c] = pb[u];
    }
  };
  load(kbeg);
  store(0);
  __syncthreads();
  int buf = 0;
  for (int k = kbeg; k < kend; k += 16) {
    const bool more = k + 16 < kend;
    if (more) load(k + 16);
    mma_quarter(acc, sB[buf], sA[buf], 16, lane, w);
    if (more) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
}

// ------------------------------------------------------------------------
// Level-batched products on the off-diagonal blocks.  item {f, c0, r0, 0}:
// the 64 x 64 tile of rows r0.. and right-hand sides c0.. of
//   dir 0:  x[rows of ABL] -= F[rows, 0:fs) z[0:fs)          (forward)
//   dir 1:  z[rows of ATL] -= F[rows, fs:n) x[fs:n)          (LU backward)
//   dir 2:  z[rows of ATL] -= F[fs:n, rows]^T x[fs:n)        (Cholesky backward)
// ------------------------------------------------------------------------
template <typename T, bool TA>
__global__ __launch_bounds__(256, 2) void k_svb_gemm(const uhmk_front* __restrict__ F,
                                                     const uhmk_item* __restrict__ items,
                                                     const T* __restrict__ arena, double* __restrict__ x,
                                                     double* __restrict__ z, int nrhs, int fwd, int ks,
                                                     double* __restrict__ part) {
  __shared__ double sA[2][16 * LD];
  __shared__ double sB[2][16 * LD];
  // blockIdx.x = item * ks + slice: with ks > 1 the K range is cut into ks
  // slices whose partial tiles go to part (reduced by k_svb_reduce)
  const int q = blockIdx.x / ks, sl = blockIdx.x - q * ks;
  const uhmk_item it = items[q];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int c0 = it.a, r0 = it.b;
  const int rend = fwd ? n : fs;
  const int kb0 = fwd ? 0 : fs, ke0 = fwd ? fs : n;
  const int len = (((ke0 - kb0) + ks - 1) / ks + 15) & ~15;
  const int kbeg = min(ke0, kb0 + sl * len), kend = min(ke0, kbeg + len);
  const double* __restrict__ S = (fwd ? z : x) + fr.xoff + (long long)c0 * n;
  d4 acc[2][2];
  zero_acc(acc);
  if (kbeg < kend)
    tile_product<T, TA>(acc, arena + fr.off, n, r0, rend, kbeg, kend, S, n, c0, nrhs, sA, sB);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (ks == 1) {
    sub_acc((fwd ? x : z) + fr.xoff, n, r0, rend, c0, nrhs, acc, lane, w);
    return;
  }
  // partial tile, layout [c][i] of the 64 x 64 tile
  double* __restrict__ P = part + (long long)blockIdx.x * 4096;
  const int cb = 32 * (w & 1), ib = 32 * (w >> 1);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int c = cb + 16 * a + 4 * g + (lane >> 4), i = ib + 16 * b + (lane & 15);
        P[c * 64 + i] = acc[a][b][g];
      }
}

// The same products for levels whose fronts are stored in FP32: operands
// rounded to FP32 on their way to LDS (the right-hand sides from FP64),
// v_mfma_f32_16x16x4f32 products and sums over 32-deep stages, the tile
// applied to the FP64 vectors.  The factor entries carry FP32 accuracy, so
// FP32 arithmetic keeps the solution at the accuracy the factor gives.
// Operand lanes as in mma_quarter; accumulator register r of lane l holds
// acc(c = 16 cb + 4 (l / 16) + r, i = 16 ib + l % 16).
typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int LDF = 80;  // row stride (floats) of the FP32 LDS panels: the four k rows a wavefront reads hit distinct banks

template <bool TA>
__global__ __launch_bounds__(256, 2) void k_svb_gemm32(const uhmk_front* __restrict__ F,
                                                       const uhmk_item* __restrict__ items,
                                                       const float* __restrict__ arena, double* __restrict__ x,
                                                       double* __restrict__ z, int nrhs, int fwd, int ks,
                                                       double* __restrict__ part) {
  constexpr int KC = 32;
  __shared__ float sA[2][KC * LDF];
  __shared__ float sB[2][KC * LDF];
  const int q = blockIdx.x / ks, sl = blockIdx.x - q * ks;
  const uhmk_item it = items[q];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int c0 = it.a, r0 = it.b;
  const int rend = fwd ? n : fs;
  const int kb0 = fwd ? 0 : fs, ke0 = fwd ? fs : n;
  const int len = (((ke0 - kb0) + ks - 1) / ks + KC - 1) / KC * KC;
  const int kbeg = min(ke0, kb0 + sl * len), kend = min(ke0, kbeg + len);
  const double* __restrict__ S = (fwd ? z : x) + fr.xoff + (long long)c0 * n;
  const float* __restrict__ A = arena + fr.off;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int cb = 32 * (w & 1), ib = 32 * (w >> 1), qq = lane & 15, kq = lane >> 4;
  f4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  float pa[8], pb[8];
  auto load = [&](int k) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (!TA) {
        const int i = tid & 63, kk = (tid >> 6) + 4 * u;
        pa[u] = (float)uhmd::ldz(A, (r0 + i) + (long long)(k + kk) * n, r0 + i < rend && k + kk < kend);
      } else {
        const int kk = tid & 31, i = (tid >> 5) + 8 * u;
        pa[u] = (float)uhmd::ldz(A, (k + kk) + (long long)(r0 + i) * n, r0 + i < rend && k + kk < kend);
      }
      const int kk = tid & 31, c = (tid >> 5) + 8 * u;
      pb[u] = (float)uhmd::ldz(S, (k + kk) + (long long)c * n, k + kk < kend && c0 + c < nrhs);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (!TA) {
        const int i = tid & 63, kk = (tid >> 6) + 4 * u;
        sA[buf][kk * LDF + i] = pa[u];
      } else {
        const int kk = tid & 31, i = (tid >> 5) + 8 * u;
        sA[buf][kk * LDF + i] = pa[u];
      }
      const int kk = tid & 31, c = (tid >> 5) + 8 * u;
      sB[buf][kk * LDF + c] = pb[u];
    }
  };
  if (kbeg < kend) {
    load(kbeg);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int k = kbeg; k < kend; k += KC) {
      const bool more = k + KC < kend;
      if (more) load(k + KC);
#pragma unroll
      for (int kk = 0; kk < KC; kk += 4) {
        const float* rb = sB[buf] + (kk + kq) * LDF + cb + qq;
        const float* ra = sA[buf] + (kk + kq) * LDF + ib + qq;
        const float b0 = rb[0], b1 = rb[16], a0 = ra[0], a1 = ra[16];
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(b0, a0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(b0, a1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(b1, a0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(b1, a1, acc[1][1], 0, 0, 0);
      }
      if (more) store(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }
  double* __restrict__ C = (fwd ? x : z) + fr.xoff;
  double* __restrict__ P = part + (long long)blockIdx.x * 4096;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int c = cb + 16 * a + 4 * kq + g, i = ib + 16 * b + qq;
        if (ks == 1) {
          if (r0 + i < rend && c0 + c < nrhs) C[(r0 + i) + (long long)(c0 + c) * n] -= (double)acc[a][b][g];
        } else {
          P[c * 64 + i] = (double)acc[a][b][g];
        }
      }
}

// dst(r0 + i, c0 + c) -= sum over the ks slices (in order) of the partial
// tiles of item q; one workgroup per item.
__global__ __launch_bounds__(256) void k_svb_reduce(const uhmk_front* __restrict__ F,
                                                    const uhmk_item* __restrict__ items, double* __restrict__ x,
                                                    double* __restrict__ z, int nrhs, int fwd, int ks,
                                                    const double* __restrict__ part) {
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int c0 = it.a, r0 = it.b;
  const int rend = fwd ? n : fs;
  double* __restrict__ D = (fwd ? x : z) + fr.xoff;
  const double* __restrict__ P = part + (long long)blockIdx.x * ks * 4096;
#pragma unroll 4
  for (int u = 0; u < 16; ++u) {
    const int e = threadIdx.x + 256 * u, c = e >> 6, i = e & 63;
    if (r0 + i >= rend || c0 + c >= nrhs) continue;
    double s = 0.0;
    for (int k = 0; k < ks; ++k) s += P[(long long)k * 4096 + e];
    D[(r0 + i) + (long long)(c0 + c) * n] -= s;
  }
}

// ------------------------------------------------------------------------
// Residual r = b - A x for all right-hand sides at once, with A the sum of
// the leaf input matrices (A = sum_e P_e K_e P_e^T, the assembly that
// DSparse_::import_element performs, UHM/src/interf/sparse.cxx:196-267):
//   leaf gather   xl_e = P_e^T x                (compact leaf layout)
//   leaf product  yl_e = K_e xl_e               (FP64 MFMA, 64 x 64 tiles)
//   row gather    r(g) = b(g) - sum of the yl entries of row g, in leaf
//                 order (deterministic), with per-column maxima of |r|,
//                 |x| and |b| for the normwise backward error.
// Compact leaf layout: leaf f's n x n_rhs block starts at loff_f * n_rhs
// (leading dimension n); item.c carries loff_f.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_leaf_gather(const uhmk_front* __restrict__ F,
                                                     const uhmk_item* __restrict__ items,
                                                     const long long* __restrict__ lg, const double* __restrict__ xg,
                                                     long long nglob, double* __restrict__ xl, int nrhs) {
  const uhmk_item it = items[blockIdx.x];
  const int n = F[it.f].n;
  const int i = it.b + threadIdx.x;
  if (i >= n) return;
  const long long g = lg[it.c + i];
  double* __restrict__ o = xl + (long long)it.c * nrhs + i;
  for (int c = 0; c < nrhs; ++c) o[(long long)c * n] = xg[g + (long long)c * nglob];
}

__global__ __launch_bounds__(256, 2) void k_leaf_mul(const uhmk_front* __restrict__ F,
                                                     const uhmk_item* __restrict__ items,
                                                     const double* __restrict__ stage,
                                                     const double* __restrict__ xl, double* __restrict__ yl,
                                                     int nrhs) {
  __shared__ double sA[2][16 * LD];
  __shared__ double sB[2][16 * LD];
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, c0 = it.a, r0 = it.b;
  const long long base = (long long)it.c * nrhs;
  d4 acc[2][2];
  zero_acc(acc);
  tile_product<double, false>(acc, stage + fr.soff, n, r0, n, 0, n, xl + base + (long long)c0 * n, n, c0, nrhs, sA,
                              sB);
  sub_acc<true>(yl + base, n, r0, n, c0, nrhs, acc, threadIdx.x & 63, threadIdx.x >> 6);
}

__device__ inline void atomic_max_nonneg(double* p, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

// grid (rows / 256, columns / RC): thread = one global row, RC columns.
constexpr int RC = 8;
__global__ __launch_bounds__(256) void k_leaf_residual(long long nrow, const long long* __restrict__ bptr,
                                                       const long long* __restrict__ lpos,
                                                       const int* __restrict__ bn, const double* __restrict__ yl,
                                                       const double* __restrict__ xg, const double* __restrict__ bg,
                                                       double* __restrict__ r, int nrhs, double* __restrict__ out) {
  __shared__ double red[3][4][RC];
  const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
  const int c0 = blockIdx.y * RC;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double mr[RC], mx[RC], mb[RC];
#pragma unroll
  for (int u = 0; u < RC; ++u) mr[u] = mx[u] = mb[u] = 0.0;
  if (g < nrow) {
    double acc[RC];
#pragma unroll
    for (int u = 0; u < RC; ++u) acc[u] = 0.0;
    for (long long q = bptr[g]; q < bptr[g + 1]; ++q) {
      const long long p = lpos[q];
      const long long ld = bn[q];
#pragma unroll
      for (int u = 0; u < RC; ++u)
        if (c0 + u < nrhs) acc[u] += yl[p + (long long)(c0 + u) * ld];
    }
#pragma unroll
    for (int u = 0; u < RC; ++u)
      if (c0 + u < nrhs) {
        const long long o = g + (long long)(c0 + u) * nrow;
        if (!bg) {
          r[o] = acc[u];  // product mode: r = A x
          continue;
        }
        const double bv = bg[o], ri = bv - acc[u];
        r[o] = ri;
        mr[u] = fabs(ri);
        mx[u] = fabs(xg[o]);
        mb[u] = fabs(bv);
      }
  }
  if (!out) return;
#pragma unroll
  for (int u = 0; u < RC; ++u) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mr[u] = fmax(mr[u], __shfl_xor(mr[u], o));
      mx[u] = fmax(mx[u], __shfl_xor(mx[u], o));
      mb[u] = fmax(mb[u], __shfl_xor(mb[u], o));
    }
  }
  if (lane == 0)
#pragma unroll
    for (int u = 0; u < RC; ++u) {
      red[0][w][u] = mr[u];
      red[1][w][u] = mx[u];
      red[2][w][u] = mb[u];
    }
  __syncthreads();
  if (threadIdx.x < 3 * RC) {
    const int k = threadIdx.x / RC, u = threadIdx.x % RC;
    if (c0 + u < nrhs) {
      const double m = fmax(fmax(red[k][0][u], red[k][1][u]), fmax(red[k][2][u], red[k][3][u]));
      // out per column: {max|r|, max row sum |A| (set by the caller), max|x|, max|b|}
      atomic_max_nonneg(out + 4 * (c0 + u) + (k == 0 ? 0 : k + 1), m);
    }
  }
}

// Global <-> front vectors over the factor rows of every front, all
// columns: dir 0 g = xf, dir 1 xf = g, dir 2 g += xf.  One thread per
// global row and 8 columns; the owning front is found by binary search in
// goff (factor rows are numbered contiguously in front order), so both
// sides are read and written in runs of consecutive rows.  Workgroup w
// takes column group w % ng of rows 256 (w / ng) ...: the column groups of a
// row range run side by side, so the workgroups in flight sweep a few
// fronts' contiguous blocks rather than a sliver of every front's block.
constexpr int GC = 8;
__global__ __launch_bounds__(256) void k_front_global_all(const uhmk_front* __restrict__ F, int nfront,
                                                          const long long* __restrict__ goff,
                                                          double* __restrict__ xf, double* __restrict__ g,
                                                          int nrhs, long long nglob, int dir, int ng) {
  const long long row = (long long)(blockIdx.x / ng) * 256 + threadIdx.x;
  if (row >= nglob) return;
  int lo = 0, hi = nfront - 1;  // largest f with goff[f] <= row
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (goff[mid] <= row) lo = mid;
    else hi = mid - 1;
  }
  const uhmk_front fr = F[lo];
  const long long xo = fr.xoff + (row - goff[lo]);
  const int c0 = (blockIdx.x % ng) * GC;
#pragma unroll
  for (int u = 0; u < GC; ++u) {
    const int c = c0 + u;
    if (c < nrhs) {
      double* __restrict__ px = xf + xo + (long long)c * fr.n;
      double* __restrict__ pg = g + row + (long long)c * nglob;
      if (dir == 0) *pg = *px;
      else if (dir == 1) *px = *pg;
      else *pg += *px;
    }
  }
}

// ------------------------------------------------------------------------
// One 64-row block step k0 of the triangular solves on ATL for one tree
// level.  item {f, c0, r, 0}: r == k0 is the block itself (y goes to z
// forward, to x backward); any other r is a 64-row chunk of ATL to update,
// below the block forward (r >= k0 + 64, rows < fs), above it backward
// (r + 64 <= k0).  MODE 0: LU, 1: Cholesky, 2: Cholesky whose diagonal
// blocks carry W = L_kk^-1 in their strict upper triangle (transposed).
// Every workgroup forms y itself, so the step needs no second launch.
// ------------------------------------------------------------------------

// Triangle of the block step into LDS for uhmd::TriStep (s[c * LD + k] =
// T(c, k) for k < c, zero elsewhere; rdiag[k] = 1 / T(k, k)).  Backward
// solves run in reversed index order (k' = 63 - k), so that TriStep's
// ascending steps walk the block from its last row up; the padding rows
// beyond kb then come first and carry rdiag 1 and zero entries.
template <typename T, int MODE, bool FWD, typename ST>
__device__ __forceinline__ void stage_tri(ST* __restrict__ s, double* __restrict__ rdiag,
                                          const T* __restrict__ B, long long n, int kb, int tid) {
  double v[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int idx = tid + 256 * u, a = idx & 63, b = idx >> 6;  // block entry (a, b), a contiguous
    bool use;
    if (FWD) use = a > b || (MODE == 1 && a == b);               // lower (L)
    else if (MODE == 0) use = a <= b;                             // upper (U)
    else use = a >= b;                                            // lower (L, read as L^T)
    v[u] = uhmd::ldz(B, a + (long long)b * n, use && a < kb && b < kb);
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int idx = tid + 256 * u, a = idx & 63, b = idx >> 6;
    // (c, k) position of entry (a, b) in the ascending solve
    int c, k;
    if (FWD) {
      c = a;
      k = b;                       // T(c, k) = L(c, k)
    } else if (MODE == 0) {
      c = 63 - a;
      k = 63 - b;                  // T'(c', k') = U(63 - c', 63 - k')
    } else {
      c = 63 - b;
      k = 63 - a;                  // T'(c', k') = L(63 - k', 63 - c')
    }
    if (c > k) s[c * LD + k] = (ST)v[u];
    else if (c < k) s[c * LD + k] = (ST)0;
    else {
      s[c * LD + k] = (ST)0;
      if (!(FWD && MODE == 0)) rdiag[k] = a < kb ? 1.0 / v[u] : 1.0;
    }
  }
}

template <typename T, int MODE, bool FWD>
__global__ __launch_bounds__(256) void k_svb_step(const uhmk_front* __restrict__ F,
                                                  const uhmk_item* __restrict__ items,
                                                  const T* __restrict__ arena, double* __restrict__ x,
                                                  double* __restrict__ z, int nrhs, int k0) {
  __shared__ double sT[64 * LD];  // triangle, then the chunk matrix (transposed)
  __shared__ double sX[64 * LD];  // x_k (k x c), then y
  __shared__ double rdiag[64];
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int c0 = it.a, r = it.b;
  const bool own = (r == k0);
  const int kb = min(64, fs - k0);
  const T* __restrict__ A = arena + fr.off;
  const T* __restrict__ Bk = A + k0 + (long long)k0 * n;
  const double* __restrict__ src = (FWD ? x : z) + fr.xoff + (long long)c0 * n;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  constexpr bool CHOL_T = (MODE != 0) && !FWD;  // chunk matrix read transposed: F(k0 + k, r + i)
  // chunk matrix into registers first (independent of y): A(i, k) -> sT[k][i]
  double pa[16];
  if (!own) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int idx = tid + 256 * u;
      if (!CHOL_T) {
        const int i = idx & 63, k = idx >> 6;
        pa[u] = uhmd::ldz(A, (r + i) + (long long)(k0 + k) * n, k < kb && (!FWD || r + i < fs));
      } else {
        const int k = idx & 63, i = idx >> 6;
        pa[u] = uhmd::ldz(A, (k0 + k) + (long long)(r + i) * n, k < kb);
      }
    }
  }
  // x_k into LDS: sX[k][c]
  {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int idx = tid + 256 * u, k = idx & 63, c = idx >> 6;
      v[u] = uhmd::ldz(src, (k0 + k) + (long long)c * n, k < kb && c0 + c < nrhs);
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int idx = tid + 256 * u, k = idx & 63, c = idx >> 6;
      sX[k * LD + c] = v[u];
    }
  }
  if constexpr (MODE == 2) {
    // sT[k][i] = Y-operator entry (i, k): forward W(i, k) (lower), backward
    // W^T(i, k) = W(k, i) (upper); W(a, b), a > b, sits at block entry
    // (b, a); the diagonal holds L(k, k), W's is its reciprocal.
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int idx = tid + 256 * u, a = idx & 63, b = idx >> 6;  // block entry (a, b)
      v[u] = uhmd::ldz(Bk, a + (long long)b * n, a <= b && b < kb);
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int idx = tid + 256 * u, a = idx & 63, b = idx >> 6;
      // entry (a, b), a < b: W(b, a); a == b: L(a, a)
      const double wv = (a == b) ? (b < kb ? 1.0 / v[u] : 0.0) : v[u];
      // forward: operator (i, k) = W(i, k): i = b, k = a  -> sT[a][b]
      // backward: operator (i, k) = W(k, i): k = b, i = a -> sT[b][a]
      // (entries with a > b are the zero triangle)
      if (FWD) sT[a * LD + b] = (a <= b) ? wv : 0.0;
      else sT[b * LD + a] = (a <= b) ? wv : 0.0;
    }
    __syncthreads();
    d4 acc[2][2];
    zero_acc(acc);
    mma_quarter(acc, sX, sT, 64, lane, w);
    __syncthreads();
    // y(i, c) = acc(c, i) -> sX[i][c]
    const int cb = 32 * (w & 1), ib = 32 * (w >> 1);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int c = cb + 16 * a + 4 * g + (lane >> 4), i = ib + 16 * b + (lane & 15);
          sX[i * LD + c] = acc[a][b][g];
        }
  } else {
    // the triangle in the front's storage type (FP32 fronts: half the LDS
    // reads of the column steps, which bound this solve; the entries are
    // exact in either type)
    typedef typename std::conditional<sizeof(T) == 4, float, double>::type ST;
    ST* __restrict__ sTs = reinterpret_cast<ST*>(sT);
    stage_tri<T, MODE, FWD>(sTs, rdiag, Bk, n, kb, tid);
    __syncthreads();
    // 16 vectors per wavefront in quad layout: lane 4 v + q holds entries
    // [16 q, 16 q + 16) of right-hand side 16 w + v (reversed backward)
    const int v = lane >> 2, q = lane & 3, c = 16 * w + v;
    double xv[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int e = 16 * q + j, k = FWD ? e : 63 - e;
      xv[j] = sX[k * LD + c];
    }
    // all 64 steps whatever kb: beyond kb the triangle holds +0, rdiag 1 and
    // x +0, so those steps leave every entry's bits as they are, and without
    // a test per step the LDS reads of later steps issue early
    if constexpr (FWD && MODE == 0) uhmd::TriStep<0, true, false, true, ST>::run(xv, sTs, LD, rdiag, 64, q);
    else uhmd::TriStep<0, false, false, true, ST>::run(xv, sTs, LD, rdiag, 64, q);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int e = 16 * q + j, k = FWD ? e : 63 - e;
      sX[k * LD + c] = xv[j];
    }
  }
  __syncthreads();
  if (own) {
    double* __restrict__ dst = (FWD ? z : x) + fr.xoff + (long long)c0 * n;
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
      const int idx = tid + 256 * u, k = idx & 63, c = idx >> 6;
      if (k < kb && c0 + c < nrhs) dst[(k0 + k) + (long long)c * n] = sX[k * LD + c];
    }
    return;
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int idx = tid + 256 * u;
    if (!CHOL_T) {
      const int i = idx & 63, k = idx >> 6;
      sT[k * LD + i] = pa[u];
    } else {
      const int k = idx & 63, i = idx >> 6;
      sT[k * LD + i] = pa[u];
    }
  }
  __syncthreads();
  d4 acc[2][2];
  zero_acc(acc);
  mma_quarter(acc, sX, sT, 64, lane, w);
  sub_acc((FWD ? x : z) + fr.xoff, n, r, FWD ? fs : k0, c0, nrhs, acc, lane, w);
}

template <typename Fn, int... I>
__device__ __forceinline__ void svb_static_for_impl(Fn&& fn, std::integer_sequence<int, I...>) {
  (fn(std::integral_constant<int, I>{}), ...);
}
// fn(integral_constant<int, i>) for i = 0 .. N-1
template <int N, typename Fn>
__device__ __forceinline__ void svb_static_for(Fn&& fn) {
  svb_static_for_impl(fn, std::make_integer_sequence<int, N>{});
}

// Block step of LU fronts with fs <= 32 at k0 = 0 (the item is the front's
// own block): one wavefront per (front, 64-column tile of right-hand
// sides); lane c solves column c0 + c in registers against the triangle
// staged in LDS step-major (sT[k][i] = T(i, k), zero outside the triangle).
// (A 64-entry version for fs <= 64 measured slower than the general step:
// one wavefront per SIMD with 4096 FMAs each.)
// Every written entry sees k_svb_step's TriStep operations in the same order
// (forward, unit L: x_i = fma(-x_k, L(i, k), x_i) for every row at steps
// k = 0 .. kb-1, zeros included; backward: x_k *= 1 / U(k, k), then the same
// with U(i, k), i < k, at steps k = kb-1 .. 0), so both give the same bits.
template <typename T, bool FWD>
__global__ __launch_bounds__(64) void k_svb_small(const uhmk_front* __restrict__ F,
                                                  const uhmk_item* __restrict__ items,
                                                  const T* __restrict__ arena, double* __restrict__ x,
                                                  double* __restrict__ z, int nrhs) {
  constexpr int KM = 32, LS = KM + 4;  // step-major rows of KM entries (+4: 16-byte aligned rows)
  // the triangle in the front's storage type (its entries are exact there;
  // each is converted to FP64 at its FMA): FP32 fronts read half the bytes
  // and hold the current and next rows in half the registers
  typedef typename std::conditional<sizeof(T) == 4, float, double>::type ST;
  __shared__ __align__(16) ST sT[KM * LS];
  __shared__ double rdg[KM];
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, kb = min(KM, fr.fs);
  const T* __restrict__ B = arena + fr.off;
  const int lane = threadIdx.x, c0 = it.a;
  for (int idx = lane; idx < KM * KM; idx += 64) {
    const int a = idx & (KM - 1), b = idx / KM;  // block entry (a, b): row a of column b
    const bool in = a < kb && b < kb && (FWD ? a > b : a <= b);
    const double v = uhmd::ldz(B, a + (long long)b * n, in);
    sT[b * LS + a] = (a == b) ? (ST)0 : (ST)v;
    if (!FWD && a == b) rdg[a] = a < kb ? 1.0 / v : 1.0;
  }
  __syncthreads();
  const bool live = c0 + lane < nrhs;
  const long long col = fr.xoff + (long long)(live ? c0 + lane : 0) * n;
  const double* __restrict__ src = (FWD ? x : z) + col;
  double xv[KM];
#pragma unroll
  for (int k = 0; k < KM; ++k) xv[k] = uhmd::ldz(src, k, live && k < kb);
  // every step's register indices at compile time (a rolled loop would
  // index xv dynamically, which puts it in scratch memory); all KM steps
  // whatever kb (beyond kb the triangle holds +0, rdg 1 and x +0: those
  // steps leave every entry's bits as they are), each step reading the next
  // step's triangle row into registers before its own FMAs.  Each step sits
  // behind a test the compiler cannot resolve (always true), which keeps the
  // steps apart: as one block, the 1024 FMAs and loads get scheduled
  // together and spill.
  int km = KM;
  asm volatile("" : "+s"(km));
  ST tc[KM], tn[KM];
#pragma unroll
  for (int i = 0; i < KM; ++i) tn[i] = sT[(FWD ? 0 : KM - 1) * LS + i];
  svb_static_for<KM>([&](auto kc) {
    constexpr int s_ = decltype(kc)::value;
    constexpr int k = FWD ? s_ : KM - 1 - s_;
    if (s_ < km) {
#pragma unroll
      for (int i = 0; i < KM; ++i) tc[i] = tn[i];
      if constexpr (s_ + 1 < KM) {
#pragma unroll
        for (int i = 0; i < KM; ++i) tn[i] = sT[(FWD ? k + 1 : k - 1) * LS + i];
      }
      if constexpr (!FWD) xv[k] = xv[k] * rdg[k];
      const double xk = xv[k];
#pragma unroll
      for (int i = 0; i < KM; ++i) xv[i] = fma(-xk, (double)tc[i], xv[i]);
    }
  });
  if (live) {
    double* __restrict__ dst = (FWD ? z : x) + col;
#pragma unroll
    for (int k = 0; k < KM; ++k)
      if (k < kb) dst[k] = xv[k];
  }
}

// out = max over rows of sum |A(row, :)| (CSR), one wavefront per row.
__global__ __launch_bounds__(256) void k_rowsum_max(long long nrow, const long long* __restrict__ rowptr,
                                                    const double* __restrict__ val, double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const long long nw = (long long)gridDim.x * 4;
  double m = 0.0;
  for (long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6); row < nrow; row += nw) {
    double a = 0.0;
    for (long long p = rowptr[row] + lane; p < rowptr[row + 1]; p += 64) a += fabs(val[p]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
    m = fmax(m, a);
  }
  if (lane == 0) atomic_max_nonneg(out, m);
}

inline int ret(hipError_t e) { return e == hipSuccess ? 0 : -(int)e; }

template <typename T>
int launch_step(int mode, int fwd, const uhmk_front* fronts, const uhmk_item* items, int n_items, int k0,
                const T* arena, double* x, double* z, int nrhs, hipStream_t stream) {
  const dim3 g(n_items), b(256);
#define UHM_SVB_STEP(M, FW) \
  hipLaunchKernelGGL((k_svb_step<T, M, FW>), g, b, 0, stream, fronts, items, arena, x, z, nrhs, k0)
  if (mode == 0) {
    if (fwd) UHM_SVB_STEP(0, true);
    else UHM_SVB_STEP(0, false);
  } else if (mode == 1) {
    if (fwd) UHM_SVB_STEP(1, true);
    else UHM_SVB_STEP(1, false);
  } else {
    if (fwd) UHM_SVB_STEP(2, true);
    else UHM_SVB_STEP(2, false);
  }
#undef UHM_SVB_STEP
  return ret(hipGetLastError());
}

template <typename T>
int launch_gemm(int dir, const uhmk_front* fronts, const uhmk_item* items, int n_items, const T* arena, double* x,
                double* z, int nrhs, int ks, double* part, hipStream_t stream) {
  const int fwd = dir == 0 ? 1 : 0;
  const dim3 g((unsigned)((long long)n_items * ks)), b(256);
  if constexpr (sizeof(T) == 4) {
    if (dir == 2)
      hipLaunchKernelGGL((k_svb_gemm32<true>), g, b, 0, stream, fronts, items, arena, x, z, nrhs, fwd, ks, part);
    else
      hipLaunchKernelGGL((k_svb_gemm32<false>), g, b, 0, stream, fronts, items, arena, x, z, nrhs, fwd, ks, part);
  } else if (dir == 2) {
    hipLaunchKernelGGL((k_svb_gemm<T, true>), g, b, 0, stream, fronts, items, arena, x, z, nrhs, fwd, ks, part);
  } else {
    hipLaunchKernelGGL((k_svb_gemm<T, false>), g, b, 0, stream, fronts, items, arena, x, z, nrhs, fwd, ks, part);
  }
  if (ks > 1)
    hipLaunchKernelGGL(k_svb_reduce, dim3(n_items), b, 0, stream, fronts, items, x, z, nrhs, fwd, ks, part);
  return ret(hipGetLastError());
}

}  // namespace

extern "C" {

int uhmk_solve_block_step(const uhmk_front* fronts, const uhmk_item* items, int n_items, int mode, int fwd, int k0,
                          const void* arena, int prec, double* x, double* z, int nrhs, hipStream_t stream) {
  if (n_items <= 0) return 0;
  if (mode < 0 || mode > 2 || k0 < 0 || k0 % 64) return -1010;
  if (prec == 0)
    return launch_step<double>(mode, fwd, fronts, items, n_items, k0, static_cast<const double*>(arena), x, z, nrhs,
                               stream);
  if (prec == 1)
    return launch_step<float>(mode, fwd, fronts, items, n_items, k0, static_cast<const float*>(arena), x, z, nrhs,
                              stream);
  return -1006;
}

int uhmk_solve_block_step_small(const uhmk_front* fronts, const uhmk_item* items, int n_items, int fwd,
                                const void* arena, int prec, double* x, double* z, int nrhs, hipStream_t stream) {
  if (n_items <= 0) return 0;
  const dim3 g(n_items), b(64);
  if (prec == 0) {
    const double* A = static_cast<const double*>(arena);
    if (fwd) hipLaunchKernelGGL((k_svb_small<double, true>), g, b, 0, stream, fronts, items, A, x, z, nrhs);
    else hipLaunchKernelGGL((k_svb_small<double, false>), g, b, 0, stream, fronts, items, A, x, z, nrhs);
  } else if (prec == 1) {
    const float* A = static_cast<const float*>(arena);
    if (fwd) hipLaunchKernelGGL((k_svb_small<float, true>), g, b, 0, stream, fronts, items, A, x, z, nrhs);
    else hipLaunchKernelGGL((k_svb_small<float, false>), g, b, 0, stream, fronts, items, A, x, z, nrhs);
  } else {
    return -1006;
  }
  return ret(hipGetLastError());
}

int uhmk_solve_block_gemm(const uhmk_front* fronts, const uhmk_item* items, int n_items, int dir,
                          const void* arena, int prec, double* x, double* z, int nrhs, int ksplit, double* part,
                          hipStream_t stream) {
  if (n_items <= 0) return 0;
  if (dir < 0 || dir > 2) return -1010;
  if (ksplit < 1 || (ksplit > 1 && !part) || (long long)n_items * ksplit > 0x7fffffffLL) return -1013;
  if (prec == 0)
    return launch_gemm<double>(dir, fronts, items, n_items, static_cast<const double*>(arena), x, z, nrhs, ksplit,
                               part, stream);
  if (prec == 1)
    return launch_gemm<float>(dir, fronts, items, n_items, static_cast<const float*>(arena), x, z, nrhs, ksplit,
                              part, stream);
  return -1006;
}

int uhmk_leaf_gather(const uhmk_front* fronts, const uhmk_item* items, int n_items, const long long* lg,
                     const double* xg, long long nglob, double* xl, int nrhs, hipStream_t stream) {
  if (n_items <= 0) return 0;
  hipLaunchKernelGGL(k_leaf_gather, dim3(n_items), dim3(256), 0, stream, fronts, items, lg, xg, nglob, xl, nrhs);
  return ret(hipGetLastError());
}

int uhmk_leaf_mul(const uhmk_front* fronts, const uhmk_item* items, int n_items, const double* stage,
                  const double* xl, double* yl, int nrhs, hipStream_t stream) {
  if (n_items <= 0) return 0;
  hipLaunchKernelGGL(k_leaf_mul, dim3(n_items), dim3(256), 0, stream, fronts, items, stage, xl, yl, nrhs);
  return ret(hipGetLastError());
}

int uhmk_leaf_residual(long long nrow, const long long* bptr, const long long* lpos, const int* bn,
                       const double* yl, const double* xg, const double* bg, double* r, int nrhs, double* out,
                       hipStream_t stream) {
  if (nrow <= 0 || nrhs <= 0) return 0;
  if ((nrhs + RC - 1) / RC > 65535) return -1011;
  const dim3 g((unsigned)((nrow + 255) / 256), (unsigned)((nrhs + RC - 1) / RC));
  hipLaunchKernelGGL(k_leaf_residual, g, dim3(256), 0, stream, nrow, bptr, lpos, bn, yl, xg, bg, r, nrhs, out);
  return ret(hipGetLastError());
}

int uhmk_rowsum_max(long long nrow, const long long* rowptr, const double* val, double* out, hipStream_t stream) {
  if (nrow <= 0) return 0;
  const long long b = (nrow + 3) / 4 < 2048 ? (nrow + 3) / 4 : 2048;
  hipLaunchKernelGGL(k_rowsum_max, dim3((unsigned)b), dim3(256), 0, stream, nrow, rowptr, val, out);
  return ret(hipGetLastError());
}

int uhmk_front_global_all(const uhmk_front* fronts, int n_fronts, const long long* goff, double* xf, double* g,
                          int nrhs, long long nglob, int dir, hipStream_t stream) {
  if (n_fronts <= 0 || nglob <= 0 || nrhs <= 0) return 0;
  if (dir < 0 || dir > 2) return -1010;
  const long long ng = (nrhs + GC - 1) / GC, nb = ng * ((nglob + 255) / 256);
  if (nb > 0x7fffffffLL) return -1011;
  hipLaunchKernelGGL(k_front_global_all, dim3((unsigned)nb), dim3(256), 0, stream, fronts, n_fronts, goff, xf, g, nrhs,
                     nglob, dir, (int)ng);
  return ret(hipGetLastError());
}

}  // extern "C"
