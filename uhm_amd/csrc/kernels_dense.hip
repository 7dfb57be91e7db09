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

#include <algorithm>
#include <type_traits>

#include "assemble_device.hpp"
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

template <int L>
__device__ __forceinline__ double readlane_d16(double v) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), L),
                          __builtin_amdgcn_readlane(__double2loint(v), L));
}
__device__ __forceinline__ double readlane_d16_rt(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

// One wavefront: C += X op(Y) for 16 x 16 blocks X, Y in LDS (row strides
// ldx, ldy), K = 16 as four v_mfma_f64_16x16x4f64; YT: op(Y) = Y^T.
// Operand lanes: X(i, k) in lane i + 16 k, op(Y)(k, j) in lane j + 16 k;
// result register r of lane l holds C(4 r + l / 16, l % 16).
template <bool YT>
__device__ __forceinline__ d4 mma16_op(const double* X, int ldx, const double* Y, int ldy, d4 c, int lane) {
#pragma unroll
  for (int k0 = 0; k0 < 16; k0 += 4) {
    const int k = k0 + (lane >> 4), q = lane & 15;
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(X[q * ldx + k], YT ? Y[q * ldy + k] : Y[k * ldy + q], c, 0, 0, 0);
  }
  return c;
}
__device__ __forceinline__ d4 mma16_ld(const double* X, int ldx, const double* Y, int ldy, d4 c, int lane) {
  return mma16_op<false>(X, ldx, Y, ldy, c, lane);
}
__device__ __forceinline__ d4 mma16(const double* X, const double* Y, int ld, d4 c, int lane) {
  return mma16_op<false>(X, ld, Y, ld, c, lane);
}

// Unblocked Cholesky of a 16-column panel by one wavefront (lane i holds
// row i of the panel in a[]).  WINV: with it W = L_BB^-1 of the panel's 16 x
// 16 diagonal block -- lane c < 16 holds column c of W in wv[] (wv = e_c on
// entry) and substitutes right-looking: at step K, W(K, c) *= 1 / L(K, K)
// (the rsqrt the step computes anyway), then W(j, c) -= L(j, K) W(K, c) for
// j > K with the L(j, K) the Cholesky update broadcasts.
// first <- K + 1 of the first non-positive pivot.
template <int K, bool WINV>
struct Chol16Step {
  __device__ __forceinline__ static void run(double (&a)[16], double (&wv)[16], int i, int& first) {
    const double d = readlane_d16<K>(a[K]);
    if (!(d > 0.0) && first == 0) first = K + 1;
    const double r = uhmd::rsqrt_d(d);
    a[K] = (i == K) ? d * r : (i > K ? a[K] * r : 0.0);
    if (WINV) wv[K] = wv[K] * r;
#pragma unroll
    for (int j = K + 1; j < 16; ++j) {
      const double ljk = readlane_d16_rt(a[K], j);
      a[j] = fma(-a[K], ljk, a[j]);
      if (WINV) wv[j] = fma(-ljk, wv[K], wv[j]);
    }
    if (WINV) {
      // W's updates happen here, at the step: the empty asm makes each entry
      // a value the step defines (sunk to the lanes that store W, they would
      // keep every broadcast L(j, K) alive to the end, spilled)
#pragma unroll
      for (int j = K; j < 16; ++j) asm volatile("" : "+v"(wv[j]));
    }
    Chol16Step<K + 1, WINV>::run(a, wv, i, first);
  }
};
template <bool WINV>
struct Chol16Step<16, WINV> {
  __device__ __forceinline__ static void run(double (&)[16], double (&)[16], int, int&) {}
};

// Cholesky of the kb x kb diagonal panel block and its inverse, one
// workgroup of four wavefronts per front, by 16-column panels.  Step B:
// wavefront 0 factors the panel rows [16 B, 64) x 16 columns, a row per lane
// (v_readlane broadcast of the column entries of the diagonal block, 1 /
// sqrt from v_rsq_f64 and two Newton steps), which gives the diagonal block
// and the blocks below it in one pass, and stores the panel's L to the
// front.  Meanwhile wavefront 1 inverts the previous panel's diagonal block
// (lane c forms column c of W(B-1, B-1)), and wavefronts 2-3 apply the
// previous panel's updates to the blocks right of panel B (deferred from
// step B - 1) and form W's block row B - 2,
// W(I, J) = -W(I, I) sum_{K=J}^{I-1} L(I, K) W(K, J).  Then panel B's
// updates of panel B + 1's blocks, A(I, B+1) -= L(I, B) L(B+1, B)^T, as
// 16 x 16 x 16 FP64 MFMA products on separate wavefronts.  The last panel
// forms its own diagonal block's inverse in its column steps, so only W's
// last two block rows follow the factorization.  Every block of L receives
// its updates in panel order.  W's strict lower part is stored transposed in
// the strict upper triangle of the diagonal block, which the Cholesky path
// never reads as matrix data (only the lower triangle of a front is used);
// the panel TRSM and the solves read it from there (its diagonal is
// 1 / L(k,k)).  Padding beyond kb is an identity block, so it never produces
// a non-positive pivot.
// LDS of the diagonal-block factorization.
struct DiagLds {
  double sL[TILE * LDW];
  double sW[TILE * LDW];
  double sS[4][16 * 17];
  int s_first;
};

// Factorization of the block held in S.sL (lower triangle of the kb x kb
// block, identity beyond kb, zeros above the diagonal; S.s_first zero, all
// visible to the workgroup; S.sW is written before it is read), then L and
// W back to the
// front at A (the block's first entry, column stride n) and the first
// non-positive pivot (c0 + its index + 1) to info[f].  Called by all 256
// threads of the workgroup.
// Store of one front entry: plain, or (COH) coherent across the XCDs'
// caches for readers inside the same launch (agent-scope relaxed atomic:
// the store carries sc1, no cache maintenance).
template <bool COH, typename T>
__device__ __forceinline__ void st_front(T* p, T v) {
  if constexpr (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *p = v;
}

template <typename T, bool COH = false>
__device__ __forceinline__ void chol_diag_lds(DiagLds& S, T* __restrict__ A, long long n, int kb, int c0,
                                              int* __restrict__ info, int f) {
  double* __restrict__ sL = S.sL;
  double* __restrict__ sW = S.sW;
  auto& sS = S.sS;
  int& s_first = S.s_first;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // 16-column blocks that hold the kb columns (the rest is identity
  // padding, which no real entry depends on)
  const int nblk = (kb + 15) >> 4, kbp = 16 * nblk;
  // L(I, J) -= L(I, P) L(J, P)^T by one wavefront
  auto tile_update = [&](int I, int J, int P) {
    d4 c = d4{0.0, 0.0, 0.0, 0.0};
    c = mma16_op<true>(sL + (16 * I) * LDW + 16 * P, LDW, sL + (16 * J) * LDW + 16 * P, LDW, c, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) sL[(16 * I + 4 * r + (lane >> 4)) * LDW + 16 * J + (lane & 15)] -= c[r];
  };
  // W(I, J) = -W(I, I) sum_{K=J}^{I-1} L(I, K) W(K, J), J < I, by one
  // wavefront (its own staging sS[w])
  auto w_block = [&](int I, int J) {
    d4 sm = d4{0.0, 0.0, 0.0, 0.0};
    for (int K = J; K < I; ++K) sm = mma16(sL + (16 * I) * LDW + 16 * K, sW + (16 * K) * LDW + 16 * J, LDW, sm, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) sS[w][(4 * r + (lane >> 4)) * 17 + (lane & 15)] = sm[r];
    // the wavefront's own LDS stores complete before its lanes read them
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    d4 wb = d4{0.0, 0.0, 0.0, 0.0};
    wb = mma16_ld(sW + (16 * I) * LDW + 16 * I, LDW, sS[w], 17, wb, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) sW[(16 * I + 4 * r + (lane >> 4)) * LDW + 16 * J + (lane & 15)] = -wb[r];
  };
  for (int B = 0; B < nblk; ++B) {
    const int b0 = 16 * B;
    if (w == 0) {
      if (lane < kbp - b0) {
        double a[16], wv[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) a[c] = (c <= lane) ? sL[(b0 + lane) * LDW + b0 + c] : 0.0;
        int first = 0;
        if (B + 1 < nblk) {
          Chol16Step<0, false>::run(a, wv, lane, first);
        } else {
#pragma unroll
          for (int c = 0; c < 16; ++c) wv[c] = (c == lane) ? 1.0 : 0.0;
          Chol16Step<0, true>::run(a, wv, lane, first);
          if (lane < 16) {
#pragma unroll
            for (int k = 0; k < 16; ++k) sW[(b0 + k) * LDW + b0 + lane] = wv[k];
          }
        }
        if (lane == 0 && first && b0 + first <= kb && s_first == 0) s_first = b0 + first;
#pragma unroll
        for (int c = 0; c < 16; ++c) sL[(b0 + lane) * LDW + b0 + c] = a[c];
        const int row = b0 + lane;
        if (row < kb) {
#pragma unroll
          for (int c = 0; c < 16; ++c)
            if (c <= lane && b0 + c < kb) st_front<COH>(A + row + (long long)(b0 + c) * n, (T)a[c]);
        }
      }
    } else if (w == 1) {
      // W(B-1, B-1): lane c forms column c; the 16 reciprocals of the
      // diagonal come first (one division per lane, in parallel), so the
      // substitution chain holds multiplications only
      if (B >= 1 && lane < 16) {
        const int p0 = b0 - 16;
        const double rdl = 1.0 / sL[(p0 + lane) * LDW + p0 + lane];
        double wc[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          double acc = (k == lane) ? 1.0 : 0.0;
#pragma unroll
          for (int m = 0; m < k; ++m) acc = fma(-sL[(p0 + k) * LDW + p0 + m], wc[m], acc);
          wc[k] = acc * readlane_d16_rt(rdl, k);
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) sW[(p0 + k) * LDW + p0 + lane] = wc[k];
      }
    } else {
      // wavefronts 2-3: panel B - 1's updates of the blocks right of panel
      // B, then W's block row B - 2
      int item = 0;
      if (B >= 1)
        for (int J = B + 1; J < nblk; ++J)
          for (int I = J; I < nblk; ++I, ++item)
            if ((item & 1) == w - 2) tile_update(I, J, B - 1);
      if (B >= 2)
        for (int J = 0; J < B - 2; ++J, ++item)
          if ((item & 1) == w - 2) w_block(B - 2, J);
    }
    __syncthreads();
    if (B + 1 < nblk) {
      // panel B's updates of panel B + 1's blocks
      for (int I = B + 1 + w; I < nblk; I += 4) tile_update(I, B + 1, B);
      __syncthreads();
    }
  }
  // W's last two block rows
  if (nblk >= 3) {
    for (int J = w; J < nblk - 2; J += 4) w_block(nblk - 2, J);
    __syncthreads();
  }
  for (int J = w; J < nblk - 1; J += 4) w_block(nblk - 1, J);
  __syncthreads();
  // the strict lower part of W, transposed, to the front
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int idx = tid + 256 * u, r = idx & 63, c = idx >> 6;
    if (c > r && c < kb) st_front<COH>(A + r + (long long)c * n, (T)sW[c * LDW + r]);
  }
  if (tid == 0 && s_first && info[f] == 0) info[f] = c0 + s_first;
}

// The kb x kb diagonal block at column c0 of front fr (kb = min(nb, fs -
// c0)) from the front into S, factored there (chol_diag_lds) and written
// back with its inverse.
template <typename T>
__device__ __forceinline__ void chol_diag_front(DiagLds& S, const uhmk_front& fr, int f, int c0, int nb,
                                                T* __restrict__ arena, int* __restrict__ info) {
  const int n = fr.n, fs = fr.fs;
  const int kb = min(nb, fs - c0);
  T* __restrict__ A = arena + fr.off + c0 + (long long)c0 * n;
  const int tid = threadIdx.x;
  {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int idx = tid + 256 * u, r = idx & 63, c = idx >> 6;
      v[u] = uhmd::ldz(A, r + (long long)c * n, r < kb && c < kb && c <= r);
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int idx = tid + 256 * u, r = idx & 63, c = idx >> 6;
      S.sL[r * LDW + c] = (r >= kb && r == c) ? 1.0 : (c <= r ? v[u] : 0.0);
    }
  }
  if (tid == 0) S.s_first = 0;
  __syncthreads();
  chol_diag_lds<T>(S, A, n, kb, c0, info, f);
}

template <typename T>
__global__ __launch_bounds__(256) void k_chol_diag(const uhmk_front* __restrict__ F,
                                                   const int* __restrict__ flist, int c0, int nb,
                                                   T* __restrict__ arena, int* __restrict__ info) {
  __shared__ DiagLds S;
  const int f = flist[blockIdx.x];
  chol_diag_front<T>(S, F[f], f, c0, nb, arena, info);
}

// Cholesky panel rows [b, b+64) below the diagonal block: X <- X L11^-T =
// X W^T with W = L11^-1 from the diagonal block (strict upper triangle:
// W transposed; diagonal: 1 / L(k,k)).  One 64 x 64 x kb FP64 MFMA product
// per workgroup; each wavefront owns a 32 x 32 quarter of the result.
constexpr int LDT = TILE + 4;

// Rows [rb, rb + 64) of front fr; sWt: LDS of TILE x LDT doubles.
template <typename T>
__device__ __forceinline__ void chol_trsm_rows(double* __restrict__ sWt, const uhmk_front& fr, int rb, int c0,
                                               int nb, T* __restrict__ arena) {
  const int n = fr.n, fs = fr.fs;
  const int kb = min(nb, fs - c0);
  T* __restrict__ A = arena + fr.off;
  // sWt[k][j] = W(j, k)
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int iw = (w & 1) * 32, jw = (w >> 1) * 32, cl = lane & 15, kq = lane >> 4;
  const T* __restrict__ D = A + c0 + (long long)c0 * n;
  double vw[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int idx = tid + 256 * u, r = idx & 63, c = idx >> 6;
    vw[u] = uhmd::ldz(D, r + (long long)c * n, r < kb && c < kb && r <= c);
  }
  // X straight into the MFMA operand registers (no LDS copy): lane l holds
  // X(iw + 16 a + l % 16, 4 s + l / 16) for k-step s, 16 consecutive rows
  // per lane group
  double xo[2][TILE / 4];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int row = rb + iw + 16 * a + cl;
#pragma unroll
    for (int ks = 0; ks < TILE / 4; ++ks) {
      const int k = 4 * ks + kq;
      xo[a][ks] = uhmd::ldz(A, row + (long long)(c0 + k) * n, row < n && k < kb);
    }
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int idx = tid + 256 * u, r = idx & 63, c = idx >> 6;
    // block entry (r, c), r <= c: W(c, r) (r < c) or L(c, c) (r == c, inverted below)
    sWt[r * LDT + c] = vw[u];
  }
  __syncthreads();
  if (tid < kb) sWt[tid * LDT + tid] = 1.0 / sWt[tid * LDT + tid];
  __syncthreads();
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < TILE / 4; ++ks) {
    const int kr = 4 * ks + kq;
    double wo[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) wo[b] = sWt[kr * LDT + jw + 16 * b + cl];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(wo[b], xo[a][ks], acc[a][b], 0, 0, 0);
  }
  // accumulator register r of lane l: D(j, i) with j = 4 r + l / 16, i = l % 16
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int row = rb + iw + 16 * a + cl;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = jw + 16 * b + 4 * r + (lane >> 4);
        if (row < n && j < kb) A[row + (long long)(c0 + j) * n] = (T)acc[a][b][r];
      }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_chol_trsm(const uhmk_front* __restrict__ F,
                                                   const uhmk_item* __restrict__ items, int c0, int nb,
                                                   T* __restrict__ arena) {
  __shared__ double sWt[TILE * LDT];
  const uhmk_item it = items[blockIdx.x];
  chol_trsm_rows<T>(sWt, F[it.f], it.b, c0, nb, arena);
}

// ------------------------------------------------------------------------
// Schur update C -= A * op(B) on T x T tiles (T = 64 or 128) on MFMA.
// A = F[i, k], B = F[k, j] (NN) or F[j, k] (NT), C = F[i, j]; K = [c0,
// min(c0+nb, fs)) for a panel update or [0, fs) for the final Schur
// complement (schur = 1).  Tiles never straddle fs, so the extent is
// min(i0 + T, fs) or min(i0 + T, n).  Four waves each own a (T/2) x (T/2)
// quarter; K advances in chunks of 16 (FP64 products) or 32 (FP32 products)
// through LDS while the next chunk is prefetched into registers.  Workgroups
// take their items through xcd_item; tiles that lie wholly inside the front
// update C without per-entry bounds checks.  TS is the storage type of the
// fronts, TM the
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

// 16 bytes of storage type T (four floats or two doubles) at any address
// aligned to T: one global_load_dwordx4; t is the aligned register type.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  static constexpr int N = 4;
  typedef float t __attribute__((ext_vector_type(4)));
  typedef float u __attribute__((ext_vector_type(4), aligned(4)));
};
template <>
struct Vec16<double> {
  static constexpr int N = 2;
  typedef double t __attribute__((ext_vector_type(2)));
  typedef double u __attribute__((ext_vector_type(2), aligned(8)));
};
// N values of type TM in registers, stored to LDS in one ds_write
template <typename TM, int N>
struct VecM {
  typedef TM t __attribute__((ext_vector_type(N)));
};

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

// DIAG (Cholesky panel updates, NT, T = 64): the workgroup of an item with
// c == 1 -- the tile (c1, c1) that is the next panel's diagonal block, c1 =
// min(c0 + nb, fs), listed first so that it starts first -- keeps its
// updated lower triangle (rounded to TS, as a store and reload would) in
// LDS and factors it there (chol_diag_lds, the k_chol_diag computation), so
// the next panel needs no diagonal launch.  Its LDS is shared with the
// product's panels, which are dead by then.
// LDS bytes of one gemm_tile: the two K-chunk panels, or the diagonal
// factorization's when that is larger (DIAG).
template <int T, typename TM, bool DIAG>
constexpr size_t gemm_lds_bytes() {
  constexpr int LD = T + 16, KC = sizeof(TM) == 4 ? 32 : 16;
  constexpr size_t panel = 2 * KC * LD * sizeof(TM);
  return (DIAG && sizeof(DiagLds) > panel) ? sizeof(DiagLds) : panel;
}

// One tile item of the update (the k_gemm computation), smem: LDS of
// gemm_lds_bytes<T, TM, DIAG>() bytes, 16-byte aligned.
// OS: storage type of the A / B operands.  With P non-null (OS = TM, NT
// products) they are read from P, a column-major copy of the update's
// operand -- rows [kend, n) of the front's columns [kbeg, kend), entry (i,
// k) at P[(i - kend) + (k - kbeg) ldp], ldp a multiple of 4 from a
// 16-byte-aligned base -- instead of the front; C is the front either way.
template <bool TRANSB, int T, typename TS, typename TM, bool DIAG, typename OS = TS>
__device__ __forceinline__ void gemm_tile(const uhmk_front& fr, const uhmk_item it, int c0, int nb, int schur,
                                          TS* __restrict__ arena, int* __restrict__ info,
                                          unsigned char* __restrict__ smem, const OS* __restrict__ P = nullptr,
                                          long long ldp = 0) {
  typedef typename Mfma<TM>::acc_t acc_t;
  constexpr int LD = T + 16;          // row stride of the LDS panels (conflict-free half-wave reads)
  constexpr int NB = T / 32;          // 16x16 blocks per wave and dimension
  constexpr int KR = 256 / T;         // k rows per load pass (one thread per i)
  // K chunk through LDS: 32 for FP32 operands (half the barriers per MFMA),
  // 16 for FP64 operands
  constexpr int KC = sizeof(TM) == 4 ? 32 : 16;
  constexpr int NL = KC / KR;         // loads per thread per operand
  constexpr int BJ = 256 / KC;        // NN B: columns per load pass
  const int n = fr.n, fs = fr.fs;
  const int kbeg = schur ? 0 : c0;
  const int kend = schur ? fs : min(c0 + nb, fs);
  const int i0 = it.a, j0 = it.b;
  const int i1 = min(i0 + T, i0 < fs ? fs : n);
  const int j1 = min(j0 + T, j0 < fs ? fs : n);
  TS* __restrict__ M = arena + fr.off;
  static_assert(!DIAG || (TRANSB && T == TILE), "fused diagonal blocks: Cholesky updates on 64-wide tiles");
  const OS* __restrict__ Po = P ? P : reinterpret_cast<const OS*>(M);
  const long long ldo = P ? ldp : (long long)n;
  const int ro = P ? kend : 0, ko = P ? kbeg : 0;
  TM* __restrict__ sA = reinterpret_cast<TM*>(smem);
  TM* __restrict__ sB = sA + KC * LD;
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
  // NN B: thread t loads k = t % KC of columns t / KC + BJ u.
  const int bk = tid % KC, bj = tid / KC;
  // A (and NT B) in 16-byte loads when storage and products have the same
  // type: VW consecutive rows per load (four floats or two doubles), thread
  // t rows VW (t % RG) .. +VW-1 of k rows t / RG + VKS u, written to LDS in
  // one ds_write (FP64 storage with FP32 products measured slower this way
  // and keeps the element loads)
  constexpr bool VEC = sizeof(OS) == sizeof(TM);
  static_assert(VEC || std::is_same<OS, TS>::value, "a separate operand copy is read with 16-byte loads");
  constexpr int VW = Vec16<OS>::N, RG = T / VW, VKS = 256 / RG, NV = KC / VKS;
  typedef typename Vec16<OS>::t vs_t;
  typedef typename Vec16<OS>::u vsu_t;
  typedef typename VecM<TM, VW>::t vm_t;
  const int vr = VW * (tid % RG), vk = tid / RG;
  vm_t va[VEC ? NV : 1], vb[VEC ? NV : 1];
  auto ldv = [&](int r0, int rend, int k) -> vm_t {
    const OS* p = Po + (r0 - ro) + (long long)(k - ko) * ldo;
    vm_t v;
    if (k < kend && r0 + VW - 1 < rend) {
      const vs_t x = *reinterpret_cast<const vsu_t*>(p);
#pragma unroll
      for (int e = 0; e < VW; ++e) v[e] = (TM)x[e];
      return v;
    }
#pragma unroll
    for (int e = 0; e < VW; ++e) v[e] = (TM)0;
    if (k < kend) {
#pragma unroll
      for (int e = 0; e < VW; ++e)
        if (r0 + e < rend) v[e] = (TM)p[e];
    }
    return v;
  };
  auto load = [&](int k0) {
    if constexpr (VEC) {
#pragma unroll
      for (int u = 0; u < NV; ++u) va[u] = ldv(i0 + vr, i1, k0 + vk + VKS * u);
      if (TRANSB) {
#pragma unroll
        for (int u = 0; u < NV; ++u) vb[u] = ldv(j0 + vr, j1, k0 + vk + VKS * u);
      }
    } else {
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int k = k0 + lk + KR * u;
        ra[u] = uhmd::ldz_as<TM>(M, (i0 + li) + (long long)k * n, i0 + li < i1 && k < kend);
      }
    }
    if (TRANSB) {
      if constexpr (VEC) return;
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int k = k0 + lk + KR * u;
        rb[u] = uhmd::ldz_as<TM>(M, (j0 + li) + (long long)k * n, j0 + li < j1 && k < kend);
      }
    } else {
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int j = bj + BJ * u;
        rb[u] = uhmd::ldz_as<TM>(M, (k0 + bk) + (long long)(j0 + j) * n, j0 + j < j1 && k0 + bk < kend);
      }
    }
  };
  if (kbeg < kend) load(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += KC) {
    if constexpr (VEC) {
#pragma unroll
      for (int u = 0; u < NV; ++u) *reinterpret_cast<vm_t*>(&sA[(vk + VKS * u) * LD + vr]) = va[u];
    } else {
#pragma unroll
      for (int u = 0; u < NL; ++u) sA[(lk + KR * u) * LD + li] = ra[u];
    }
    if (TRANSB) {
      if constexpr (VEC) {
#pragma unroll
        for (int u = 0; u < NV; ++u) *reinterpret_cast<vm_t*>(&sB[(vk + VKS * u) * LD + vr]) = vb[u];
      } else {
#pragma unroll
        for (int u = 0; u < NL; ++u) sB[(lk + KR * u) * LD + li] = rb[u];
      }
    } else {
#pragma unroll
      for (int u = 0; u < NL; ++u) sB[bk * LD + bj + BJ * u] = rb[u];
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
  if (DIAG && it.c == 1) {
    // the next panel's diagonal block: tile (c1, c1), kb = i1 - i0 rows
    DiagLds& S = *reinterpret_cast<DiagLds*>(smem);
    const int kb = i1 - i0;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int idx = tid + 256 * u, r = idx & 63, c = idx >> 6;
      S.sL[r * LDW + c] = (r >= kb && r == c) ? 1.0 : 0.0;
    }
    if (tid == 0) S.s_first = 0;
    __syncthreads();
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int r = iw + a * 16 + (lane & 15);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int c = jw + b * 16 + Mfma<TM>::col(q, lane);
          if (r < kb && c <= r) {
            const TS* e = M + (i0 + r) + (long long)(j0 + c) * n;
            S.sL[r * LDW + c] = (double)(TS)((double)*e - (double)acc[a][b][q]);
          }
        }
      }
    __syncthreads();
    chol_diag_lds<TS>(S, M + i0 + (long long)j0 * n, n, kb, i0, info, it.f);
    return;
  }
  if (i1 - i0 == T && j1 - j0 == T) {
    // whole tile inside the front: no per-entry bounds checks, so the
    // read-modify-writes issue back to back
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int i = i0 + iw + a * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int j = j0 + jw + b * 16 + Mfma<TM>::col(r, lane);
          TS* c = M + i + (long long)j * n;
          *c = (TS)((double)*c - (double)acc[a][b][r]);
        }
      }
    return;
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

template <bool TRANSB, int T, typename TS, typename TM, bool DIAG = false>
__global__ __launch_bounds__(256, 2) void k_gemm(const uhmk_front* __restrict__ F,
                                              const uhmk_item* __restrict__ items, int c0, int nb, int schur,
                                              TS* __restrict__ arena, int* __restrict__ info = nullptr) {
  __shared__ __align__(16) unsigned char smem[gemm_lds_bytes<T, TM, DIAG>()];
  const uhmk_item it = items[uhmd::xcd_item(blockIdx.x, gridDim.x)];
  gemm_tile<TRANSB, T, TS, TM, DIAG>(F[it.f], it, c0, nb, schur, arena, info, smem);
}

template <typename TS, typename TM>
int launch_gemm_diag(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, TS* arena,
                     int* info, hipStream_t stream) {
  hipLaunchKernelGGL((k_gemm<true, 64, TS, TM, true>), dim3(n_items), dim3(256), 0, stream, fronts, items, c0, nb, 0,
                     arena, info);
  return 0;
}

// ------------------------------------------------------------------------
// The same update with BF16 operands and FP32 accumulation
// (v_mfma_f32_16x16x32_bf16): A and B are rounded once to BF16 (nearest
// even) on their way to LDS, the products and sums are FP32, and C is updated in
// its storage type.  Both LDS panels are stored [row][k] with k contiguous,
// so that a lane fetches its eight consecutive k of one row with one
// ds_read_b128; rows are LDK = 40 BF16 (80 bytes) apart, which spreads the
// sixteen rows a quarter-wave reads over all 64 banks.  One K chunk of 32 is
// one MFMA k-step.
// Operand layout: lane l supplies row l % 16 and k = 8 (l / 16) .. +7 of both
// operands (any k order common to the two operands gives the same sum);
// accumulator register r of lane l holds D[4 (l / 16) + r][l % 16].  As in
// k_gemm the product is formed transposed (D = B^T A^T), so a 16-lane group
// stores 16 consecutive rows of C.
// Staging: A (and NT B, whose rows are also contiguous in memory) -- thread
// t loads row t % T at KPT consecutive k and writes them packed; NN B
// (columns contiguous in memory) -- thread t loads the k pair 2 (t % 16),
// +1 of columns t / 16 + 16 u and writes it as one dword.
// ------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// One rounding to BF16, nearest even, from either storage type (an FP64
// value is not rounded to FP32 first).
template <typename TS>
__device__ __forceinline__ __bf16 to_bf16(TS v) {
  return (__bf16)v;
}

template <bool TRANSB, int T, typename TS>
__global__ __launch_bounds__(256, 2) void k_gemm_bf16(const uhmk_front* __restrict__ F,
                                                   const uhmk_item* __restrict__ items, int c0, int nb, int schur,
                                                   TS* __restrict__ arena) {
  constexpr int KC = 32;             // k per chunk: one MFMA k-step
  constexpr int LDK = KC + 8;        // BF16 row stride of the LDS panels
  constexpr int NB = T / 32;         // 16x16 blocks per wave and dimension
  constexpr int KPT = KC * T / 256;  // A / NT B: consecutive k per thread (16 or 8)
  constexpr int JP = T / 16;         // NN B: columns per thread
  static_assert(KPT % 8 == 0, "A rows are written in groups of eight");
  const uhmk_item it = items[uhmd::xcd_item(blockIdx.x, gridDim.x)];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int kbeg = schur ? 0 : c0;
  const int kend = schur ? fs : min(c0 + nb, fs);
  const int i0 = it.a, j0 = it.b;
  const int i1 = min(i0 + T, i0 < fs ? fs : n);
  const int j1 = min(j0 + T, j0 < fs ? fs : n);
  TS* __restrict__ M = arena + fr.off;
  __shared__ __align__(16) __bf16 sA[T * LDK];
  __shared__ __align__(16) __bf16 sB[T * LDK];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int iw = (w & 1) * (T / 2), jw = (w >> 1) * (T / 2);

  f4 acc[NB][NB];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = f4{0, 0, 0, 0};

  TS ra[KPT], rb[KPT];
  const int li = tid % T, lk = (tid / T) * KPT;    // A / NT B: row, first k
  const int bp = 2 * (tid % 16), bj = tid / 16;     // NN B: k pair, first column
  const bool a_ok = i0 + li < i1, bt_ok = j0 + li < j1;
  auto load = [&](int k0) {
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
      const int k = k0 + lk + u;
      ra[u] = uhmd::ldz_as<TS>(M, (i0 + li) + (long long)k * n, a_ok && k < kend);
    }
    if (TRANSB) {
#pragma unroll
      for (int u = 0; u < KPT; ++u) {
        const int k = k0 + lk + u;
        rb[u] = uhmd::ldz_as<TS>(M, (j0 + li) + (long long)k * n, bt_ok && k < kend);
      }
    } else {
#pragma unroll
      for (int u = 0; u < JP; ++u) {
        const int j = j0 + bj + 16 * u;
        const long long base = (long long)j * n + k0 + bp;
        rb[2 * u] = uhmd::ldz_as<TS>(M, base, j < j1 && k0 + bp < kend);
        rb[2 * u + 1] = uhmd::ldz_as<TS>(M, base + 1, j < j1 && k0 + bp + 1 < kend);
      }
    }
  };
  if (kbeg < kend) load(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += KC) {
#pragma unroll
    for (int h = 0; h < KPT / 8; ++h) {
      bf16x8 v;
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = to_bf16(ra[8 * h + e]);
      *reinterpret_cast<bf16x8*>(&sA[li * LDK + lk + 8 * h]) = v;
    }
    if (TRANSB) {
#pragma unroll
      for (int h = 0; h < KPT / 8; ++h) {
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = to_bf16(rb[8 * h + e]);
        *reinterpret_cast<bf16x8*>(&sB[li * LDK + lk + 8 * h]) = v;
      }
    } else {
#pragma unroll
      for (int u = 0; u < JP; ++u) {
        bf16x2 v;
        v[0] = to_bf16(rb[2 * u]);
        v[1] = to_bf16(rb[2 * u + 1]);
        *reinterpret_cast<bf16x2*>(&sB[(bj + 16 * u) * LDK + bp]) = v;
      }
    }
    __syncthreads();
    if (k0 + KC < kend) load(k0 + KC);
    {
      const int kq = 8 * (lane >> 4), cl = lane & 15;
      bf16x8 bo[NB], ao[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) bo[b] = *reinterpret_cast<const bf16x8*>(&sB[(jw + b * 16 + cl) * LDK + kq]);
#pragma unroll
      for (int a = 0; a < NB; ++a) ao[a] = *reinterpret_cast<const bf16x8*>(&sA[(iw + a * 16 + cl) * LDK + kq]);
#pragma unroll
      for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bo[b], ao[a], acc[a][b], 0, 0, 0);
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
        const int j = j0 + jw + b * 16 + 4 * (lane >> 4) + r;
        if (i < i1 && j < j1) {
          TS* c = M + i + (long long)j * n;
          *c = (TS)((double)*c - (double)acc[a][b][r]);
        }
      }
    }
}

template <typename TS>
int launch_gemm_bf16(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, int schur,
                     int transb, int tile, TS* arena, hipStream_t stream) {
  dim3 g(n_items), b(256);
  if (tile == 128) {
    if (transb) hipLaunchKernelGGL((k_gemm_bf16<true, 128, TS>), g, b, 0, stream, fronts, items, c0, nb, schur, arena);
    else hipLaunchKernelGGL((k_gemm_bf16<false, 128, TS>), g, b, 0, stream, fronts, items, c0, nb, schur, arena);
  } else if (tile == 64) {
    if (transb) hipLaunchKernelGGL((k_gemm_bf16<true, 64, TS>), g, b, 0, stream, fronts, items, c0, nb, schur, arena);
    else hipLaunchKernelGGL((k_gemm_bf16<false, 64, TS>), g, b, 0, stream, fronts, items, c0, nb, schur, arena);
  } else {
    return -1004;
  }
  return 0;
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

// ------------------------------------------------------------------------
// Cholesky panel group as one persistent launch: the columns [g0, ge) of
// every front of a launch, ge = min(g0 + gw, fs), factored left-looking
// inside the group (the chol.cxx:60-92 steps -- diagonal block, panel
// TRSM, update -- for panels of 64 columns, LINAL/src/dense/chol.cxx:67,
// 80-84, 86-90; the updates from columns left of g0 came from earlier
// trailing updates).  One workgroup per 64-row tile r of a front's rows
// [g0, n) (tiles cut at fs; item {f, r, first row, flag slot}); tile r
// runs the panels p = 0 .. (r < P ? r : P - 1) in order, P = panels of the
// group:
//   C = A(r, p) - sum_{q < p} L(r, q) L(p, q)^T    (MFMA, K = 64 p)
//   r == p: C = L(p, p) L(p, p)^T and W_p = L(p, p)^-1 (chol_diag_lds), to
//           the front; publish
//   r != p: wait until tile p published W_p; L(r, p) = C W_p^T (FP64
//           MFMA); tiles r < P publish.
// Tile r < P publishes its progress (panels done) in flags[slot + r];
// waiters spin on it (one thread, bounded: after SPIN_LIMIT polls flags[0]
// is raised and the launch finishes with wrong values instead of hanging).
// The factor entries written inside the launch (L and W) are stored and
// read as agent-scope relaxed atomics (sc1: coherent across the XCDs'
// L2 caches, without the whole-cache write-back / invalidation that
// acquire / release fences at agent scope cost), and a tile's stores have
// completed (s_waitcnt) before it publishes.  A tile waits only on tiles of
// the same front with a smaller index, listed earlier in the launch, so the
// in-order dispatch of workgroups always leaves the awaited tile running or
// done.  The chain per panel is diagonal block -> TRSM of the next tile ->
// its rank-64 diagonal update -> next diagonal block, without launches in
// between.  TS storage, TM MFMA type of the updates (as k_gemm); the
// diagonal block and the TRSM run in FP64.
// ------------------------------------------------------------------------
// Workgroup barrier that orders LDS only: global loads in flight stay in
// flight across it (__syncthreads waits for them)
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

constexpr int GROUP_MAX_PANELS = 16;
constexpr int SPIN_LIMIT = 1 << 22;

__device__ __forceinline__ void group_wait(int* flag, int target, int* err) {
  if (threadIdx.x == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
    int spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > SPIN_LIMIT) {
        atomicAdd(err, 1);
        break;
      }
    }
  }
  __syncthreads();
}

// every thread's coherent stores of this step complete, then (flag
// non-null) the progress value published by one thread
__device__ __forceinline__ void group_publish(int* flag, int v) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (flag && threadIdx.x == 0) __hip_atomic_store(flag, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Coherent load (agent-scope relaxed atomic: sc1) of data that another
// workgroup of the same launch may have stored, zero when !ok.
template <typename TM, typename T>
__device__ __forceinline__ TM ld_coh(const T* p, long long idx, bool ok) {
  return ok ? (TM)__hip_atomic_load(const_cast<T*>(p) + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : (TM)0;
}

template <typename TS, typename TM>
__global__ __launch_bounds__(256, 2) void k_chol_group(const uhmk_front* __restrict__ F,
                                                    const uhmk_item* __restrict__ items, int g0, int gw,
                                                    TS* __restrict__ arena, int* __restrict__ flags,
                                                    int* __restrict__ info) {
  typedef typename Mfma<TM>::acc_t acc_t;
  constexpr int T = TILE, LD = T + 16, NB = 2, KR = 256 / T;
  constexpr int KC = sizeof(TM) == 4 ? 32 : 16;
  constexpr int NL = KC / KR;
  // the update staging stays below sWt, which lives through the tiles of a panel step
  static_assert((size_t)2 * KC * LD * sizeof(TM) <= (size_t)TILE * LDW * sizeof(double), "staging below sWt");
  static_assert(sizeof(DiagLds) >= T * LD * sizeof(TM), "the L(r, p) operand fits the diagonal LDS");
  static_assert(sizeof(DiagLds) >= (size_t)(TILE * LDW + TILE * LDT) * sizeof(double), "TRSM fits");
  __shared__ __align__(16) unsigned char smem[sizeof(DiagLds)];
  TM* __restrict__ sA = reinterpret_cast<TM*>(smem);
  TM* __restrict__ sB = sA + KC * LD;
  double* __restrict__ sC = reinterpret_cast<double*>(smem);  // 64 x LDW, row major (= DiagLds::sL)
  double* __restrict__ sWt = sC + TILE * LDW;                 // TRSM: sWt[k][j] = W(j, k)
  DiagLds& S = *reinterpret_cast<DiagLds*>(smem);

  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int ge = min(g0 + gw, fs);
  const int P = (ge - g0 + T - 1) / T;
  // row tiles of [g0, n) cut at fs: the item's tiles r .. r + ntile - 1
  const int ratl = (fs - g0 + T - 1) / T;
  auto tile_row = [&](int t) { return t < ratl ? g0 + T * t : fs + T * (t - ratl); };
  const int r = it.a, ntile = it.b;
  const bool diag_tile = r < P;
  int* err = flags;
  int* fl = flags + it.c;
  TS* __restrict__ M = arena + fr.off;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int iw = (w & 1) * 32, jw = (w >> 1) * 32, cl = lane & 15, kq = lane >> 4;
  const int li = tid % T, lk = tid / T;

  // acc = L(rows [i0, i0 + h), [g0, pc0)) L(p, [g0, pc0))^T for panel p
  auto update = [&](int i0, int h, int pc0, int kb, acc_t (&acc)[NB][NB]) {
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
    TM ra[NL], rb[NL];
    auto load = [&](int k0) {
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int k = k0 + lk + KR * u;
        ra[u] = ld_coh<TM>(M, (i0 + li) + (long long)k * n, li < h && k < pc0);
        rb[u] = ld_coh<TM>(M, (pc0 + li) + (long long)k * n, li < kb && k < pc0);
      }
    };
    if (g0 < pc0) load(g0);
    for (int k0 = g0; k0 < pc0; k0 += KC) {
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        sA[(lk + KR * u) * LD + li] = ra[u];
        sB[(lk + KR * u) * LD + li] = rb[u];
      }
      __syncthreads();
      if (k0 + KC < pc0) load(k0 + KC);
#pragma unroll
      for (int kk = 0; kk < KC; kk += 4) {
        const int kr = kk + kq;
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
  };

  // The panel steps of the item's tiles; DIAGT: a tile of the group's
  // diagonal range (one tile, ends with its own diagonal block).  Two
  // instances, so that the accumulator of the diagonal block is live only
  // where it is used.
  auto run = [&](auto diag_c) {
    constexpr bool DIAGT = decltype(diag_c)::value;
    // panels whose diagonal block belongs to another tile; a tile of the
    // group's diagonal range also accumulates sum_p L(r, p) L(r, p)^T for its
    // own diagonal block as each L(r, p) appears, so that block needs no
    // update of its own once the last TRSM is done
    const int nsolve = DIAGT ? r : P;
    // the own diagonal block's entries as the earlier launches left them
    // (nothing else in this launch writes them): loaded before the last
    // wait for W, so the loads complete while the tile waits
    TS mv[NB][NB][4];
    auto load_mv = [&]() {
      const int dc0 = g0 + T * r, dkb = min(T, ge - dc0);
#pragma unroll
      for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const int ii = iw + a * 16 + cl;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int jj = jw + b * 16 + Mfma<TM>::col(q, lane);
            mv[a][b][q] = M[ii < dkb && jj <= ii ? (dc0 + ii) + (long long)(dc0 + jj) * n : dc0 + (long long)dc0 * n];
          }
        }
    };
    acc_t accd[NB][NB];
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) accd[a][b] = acc_t{0, 0, 0, 0};
    for (int p = 0; p < nsolve; ++p) {
      const int pc0 = g0 + T * p, kb = min(T, ge - pc0);
      // L(p, q < p) of tile p done (own rows: earlier steps of this workgroup)
      if (p > 0) group_wait(fl + p, p, err);
      for (int t = 0; t < ntile; ++t) {
        const int i0 = tile_row(r + t);
        const int h = min(i0 + T, i0 < fs ? fs : n) - i0;
        acc_t acc[NB][NB];
        update(i0, h, pc0, kb, acc);
        // C (rounded to TS, as a store and reload would) into LDS, zero
        // outside h x kb
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            const int ii = iw + a * 16 + cl;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int jj = jw + b * 16 + Mfma<TM>::col(q, lane);
              double v = 0.0;
              if (ii < h && jj < kb) {
                const double c = (double)M[(i0 + ii) + (long long)(pc0 + jj) * n];
                v = (double)(TS)(c - (double)acc[a][b][q]);
              }
              sC[ii * LDW + jj] = v;
            }
          }
        // W_p from the diagonal block of panel p (strict upper: W transposed;
        // diagonal: L(k, k), inverted below), once per panel step
        if (t > 0) {
          __syncthreads();
        } else {
          group_wait(fl + p, p + 1, err);
          const TS* __restrict__ D = M + pc0 + (long long)pc0 * n;
#pragma unroll 8
          for (int u = 0; u < 16; ++u) {
            const int idx = tid + 256 * u, rr = idx & 63, cc = idx >> 6;
            sWt[rr * LDT + cc] = ld_coh<double>(D, rr + (long long)cc * n, rr < kb && cc < kb && rr <= cc);
          }
          __syncthreads();
          if (tid < kb) sWt[tid * LDT + tid] = 1.0 / sWt[tid * LDT + tid];
          __syncthreads();
        }
        // L(r, p) = C W^T: one 64 x 64 x 64 FP64 MFMA product, wavefront w owns
        // the 32 x 32 quarter (iw, jw) (as k_chol_trsm)
        d4 acc2[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) acc2[a][b] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int ks = 0; ks < TILE / 4; ++ks) {
          const int kr = 4 * ks + kq;
          double wo[2], xo[2];
#pragma unroll
          for (int b = 0; b < 2; ++b) wo[b] = sWt[kr * LDT + jw + 16 * b + cl];
#pragma unroll
          for (int a = 0; a < 2; ++a) xo[a] = sC[(iw + 16 * a + cl) * LDW + kr];
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc2[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(wo[b], xo[a], acc2[a][b], 0, 0, 0);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            const int ii = iw + 16 * a + cl;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int jj = jw + 16 * b + 4 * q + kq;
              if (ii < h && jj < kb) st_front<true>(M + (i0 + ii) + (long long)(pc0 + jj) * n, (TS)acc2[a][b][q]);
            }
          }
        if (diag_tile) {
          // accd += X X^T with X = L(r, p) as stored (TS), in TM products
          __syncthreads();  // sC / sWt read by every wavefront
          TM* __restrict__ sX = reinterpret_cast<TM*>(smem);  // sX[k][i] = X(i, k)
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
              const int ii = iw + 16 * a + cl;
#pragma unroll
              for (int q = 0; q < 4; ++q) {
                const int jj = jw + 16 * b + 4 * q + kq;
                sX[jj * LD + ii] = (ii < h && jj < kb) ? (TM)(TS)acc2[a][b][q] : (TM)0;
              }
            }
          __syncthreads();
#pragma unroll
          for (int kk = 0; kk < T; kk += 4) {
            const int kr = kk + kq;
            TM bo[NB], ao[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) bo[b] = sX[kr * LD + jw + b * 16 + cl];
#pragma unroll
            for (int a = 0; a < NB; ++a) ao[a] = sX[kr * LD + iw + a * 16 + cl];
#pragma unroll
            for (int a = 0; a < NB; ++a)
#pragma unroll
              for (int b = 0; b < NB; ++b) accd[a][b] = Mfma<TM>::run(bo[b], ao[a], accd[a][b]);
          }
        }
        if (t + 1 < ntile) __syncthreads();  // sC read before the next tile's staging
      }
      // own rows visible to this workgroup's next steps; tiles of the
      // group's diagonal range also tell the others
      group_publish(DIAGT ? fl + r : nullptr, p + 1);
    }
    if constexpr (DIAGT) {

    // the diagonal block of panel r (rows = panel columns) minus the
    // accumulated products, its lower triangle (rounded to TS as a store and
    // reload would, identity beyond kb) through chol_diag_lds
    {
      const int pc0 = g0 + T * r, kb = min(T, ge - pc0);
      // the loads stay in flight across the LDS set-up
      load_mv();
      lds_sync();  // sX read by every wavefront
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int idx = tid + 256 * u, rr = idx & 63, cc = idx >> 6;
        S.sL[rr * LDW + cc] = (rr >= kb && rr == cc) ? 1.0 : 0.0;
      }
      if (tid == 0) S.s_first = 0;
      lds_sync();
#pragma unroll
      for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const int ii = iw + a * 16 + cl;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int jj = jw + b * 16 + Mfma<TM>::col(q, lane);
            if (ii < kb && jj <= ii) S.sL[ii * LDW + jj] = (double)(TS)((double)mv[a][b][q] - (double)accd[a][b][q]);
          }
        }
      lds_sync();
      chol_diag_lds<TS, true>(S, M + pc0 + (long long)pc0 * n, n, kb, pc0, info, it.f);
      group_publish(fl + r, r + 1);
    }
    }
  };
  if (diag_tile) run(std::true_type{});
  else run(std::false_type{});
}

// ------------------------------------------------------------------------
// Whole small fronts in one launch (Cholesky, fs <= 64): one workgroup per
// front runs, in order, the computations of the assembly, diagonal, panel
// TRSM and Schur complement launches of a tree level on its own front --
// extend-add of the lower 64 x 64 tiles (uhmd::asm_tile), the fs x fs
// diagonal block and its inverse (chol_diag_front), the rows below it
// (chol_trsm_rows) and ABR -= ABL ABL^T on 64-wide tiles (gemm_tile,
// K = [0, fs), TM products) -- so a level of many small fronts takes one
// launch instead of four, the front stays in the L2 between the steps, and
// the factors are the same bits as the four launches give (the same
// operations in the same order).  The reference steps: op_merge then
// chol_hier on each element, leaves first (UHM/src/operation/element.cxx:
// 200-241, UHM/src/matrix/uhm/fla/chol/decompose.cxx:85-98).  A front with
// fs > 64 is not touched and reports info[f] = -1.
// ------------------------------------------------------------------------
template <typename TS, typename TM>
__global__ __launch_bounds__(256, 2) void k_chol_front_small(const uhmk_front* __restrict__ F,
                                                          const int* __restrict__ flist,
                                                          const int* __restrict__ src_beg,
                                                          const uhmk_source* __restrict__ srcs,
                                                          const int* __restrict__ inv, double* __restrict__ arena64,
                                                          float* __restrict__ arena32,
                                                          const double* __restrict__ stage,
                                                          int* __restrict__ info) {
  __shared__ __align__(16) unsigned char smem[sizeof(DiagLds)];
  static_assert(sizeof(DiagLds) >= TILE * LDT * sizeof(double), "TRSM staging fits");
  static_assert(sizeof(DiagLds) >= gemm_lds_bytes<64, TM, false>(), "product panels fit");
  static_assert(sizeof(DiagLds) >= sizeof(uhmd::AsmSources), "source table fits");
  const int f = flist[blockIdx.x];
  const uhmk_front fr = F[f];
  const int n = fr.n, fs = fr.fs;
  if (fs > TILE) {
    if (threadIdx.x == 0) info[f] = -1;
    return;
  }
  TS* __restrict__ arena;
  if constexpr (sizeof(TS) == 8) arena = reinterpret_cast<TS*>(arena64);
  else arena = reinterpret_cast<TS*>(arena32);
  {
    uhmd::AsmSources& S = *reinterpret_cast<uhmd::AsmSources*>(smem);
    uhmd::asm_sources(S, F, f, src_beg, srcs);
    __syncthreads();
    for (int j0 = 0; j0 < n; j0 += TILE)
      for (int i0 = j0; i0 < n; i0 += TILE) uhmd::asm_tile<TS>(S, n, fr.off, i0, j0, inv, arena64, arena32, stage);
    __syncthreads();
  }
  if (fs == 0) return;
  chol_diag_front<TS>(*reinterpret_cast<DiagLds*>(smem), fr, f, 0, TILE, arena, info);
  __syncthreads();
  for (int rb = fs; rb < n; rb += TILE) {
    chol_trsm_rows<TS>(reinterpret_cast<double*>(smem), fr, rb, 0, TILE, arena);
    __syncthreads();
  }
  for (int j0 = fs; j0 < n; j0 += TILE)
    for (int i0 = j0; i0 < n; i0 += TILE)
      gemm_tile<true, TILE, TS, TM, false>(fr, uhmk_item{f, i0, j0, 0}, 0, TILE, 1, arena, nullptr, smem);
}

// ------------------------------------------------------------------------
// Cholesky updates of FP64 fronts with FP32 MFMA on an FP32 copy of their
// operand: the trailing updates C -= L(:, [c0, c1)) L(:, [c0, c1))^T of the
// rows [c1, n), c1 = min(c0 + nb, fs), and (schur) the final Schur
// complement ABR -= ABL ABL^T (c0 = 0, c1 = fs; the reference's gemm /
// syrk, LINAL/src/dense/chol.cxx:86-90, UHM/src/matrix/uhm/fla/chol/
// decompose.cxx:95-96).  k_p32_pack rounds the operand to FP32 once into a
// column-major copy (ldp = n - c1 rounded up to 4, padding rows zero;
// items {f, 64-row tile, 64-column tile, offset / 64}); k_gemm_p32 then runs
// the tiles of gemm_tile on it with 16-byte loads (items {f, i0, j0, offset
// / 64}).  The operands reach LDS as the same FP32 values as the element
// loads of the front give (one rounding of each entry), so C is the same
// bits as k_gemm's on the front itself.
// ------------------------------------------------------------------------
__host__ __device__ inline long long p32_ld(int n, int c1) { return ((long long)(n - c1) + 3) & ~3LL; }

template <typename TS>
__global__ __launch_bounds__(256) void k_p32_pack(const uhmk_front* __restrict__ F,
                                                  const uhmk_item* __restrict__ items, int c0, int nb, int schur,
                                                  const TS* __restrict__ arena, float* __restrict__ P) {
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int kbeg = schur ? 0 : c0, kend = schur ? fs : min(c0 + nb, fs);
  const int rows = n - kend;
  const long long ldp = p32_ld(n, kend);
  const TS* __restrict__ M = arena + fr.off;
  float* __restrict__ D = P + 64LL * it.c;
  const int tid = threadIdx.x, r = tid & 63;
  const long long row = 64LL * it.a + r;
  float v[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int k = kbeg + 64 * it.b + (tid >> 6) + 4 * u;
    v[u] = (row < rows && k < kend) ? (float)M[(kend + row) + (long long)k * n] : 0.0f;
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int k = kbeg + 64 * it.b + (tid >> 6) + 4 * u;
    if (row < ldp && k < kend) D[row + (long long)(k - kbeg) * ldp] = v[u];
  }
}

template <typename TS>
__global__ __launch_bounds__(256, 2) void k_gemm_p32(const uhmk_front* __restrict__ F,
                                                  const uhmk_item* __restrict__ items, int c0, int nb, int schur,
                                                  TS* __restrict__ arena, const float* __restrict__ P) {
  __shared__ __align__(16) unsigned char smem[gemm_lds_bytes<64, float, false>()];
  const uhmk_item it = items[uhmd::xcd_item(blockIdx.x, gridDim.x)];
  const uhmk_front fr = F[it.f];
  const int kend = schur ? fr.fs : min(c0 + nb, fr.fs);
  gemm_tile<true, 64, TS, float, false, float>(fr, it, c0, nb, schur, arena, nullptr, smem, P + 64LL * it.c,
                                               p32_ld(fr.n, kend));
}

template <typename TS, typename TM>
int launch_chol_group(const uhmk_front* fronts, const uhmk_item* items, int n_items, int g0, int gw, TS* arena,
                      int* flags, int* info, hipStream_t stream) {
  hipLaunchKernelGGL((k_chol_group<TS, TM>), dim3(n_items), dim3(256), 0, stream, fronts, items, g0, gw, arena, flags,
                     info);
  return 0;
}

inline int ret(hipError_t e) { return e == hipSuccess ? 0 : -(int)e; }

}  // namespace

extern "C" {

int uhmk_chol_group(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int ncol, void* arena,
                    int prec, int mma, int* flags, int* info, hipStream_t stream) {
  if (n_items <= 0) return 0;
  if (ncol < TILE || ncol % TILE || ncol > TILE * GROUP_MAX_PANELS) return -1000;
  int rc;
  if (prec == 0 && mma == 0)
    rc = launch_chol_group<double, double>(fronts, items, n_items, c0, ncol, static_cast<double*>(arena), flags, info,
                                           stream);
  else if (prec == 0 && mma == 1)
    rc = launch_chol_group<double, float>(fronts, items, n_items, c0, ncol, static_cast<double*>(arena), flags, info,
                                          stream);
  else if (prec == 1 && mma == 1)
    rc = launch_chol_group<float, float>(fronts, items, n_items, c0, ncol, static_cast<float*>(arena), flags, info,
                                         stream);
  else
    return -1006;
  if (rc) return rc;
  return ret(hipGetLastError());
}

int uhmk_chol_front_small(const uhmk_front* fronts, const int* flist, int n_fronts, const int* src_beg,
                          const uhmk_source* srcs, const int* inv, double* arena64, float* arena32, int prec, int mma,
                          const double* stage, int* info, hipStream_t stream) {
  if (n_fronts <= 0) return 0;
  const dim3 g(n_fronts), b(256);
  if (prec == 0 && mma == 0)
    hipLaunchKernelGGL((k_chol_front_small<double, double>), g, b, 0, stream, fronts, flist, src_beg, srcs, inv,
                       arena64, arena32, stage, info);
  else if (prec == 0 && mma == 1)
    hipLaunchKernelGGL((k_chol_front_small<double, float>), g, b, 0, stream, fronts, flist, src_beg, srcs, inv,
                       arena64, arena32, stage, info);
  else if (prec == 1 && mma == 1)
    hipLaunchKernelGGL((k_chol_front_small<float, float>), g, b, 0, stream, fronts, flist, src_beg, srcs, inv,
                       arena64, arena32, stage, info);
  else
    return -1006;
  return ret(hipGetLastError());
}

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

int uhmk_chol_gemm_diag(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, void* arena,
                        int prec, int mma, int* info, hipStream_t stream) {
  if (n_items <= 0) return 0;
  if (nb != TILE) return -1000;
  int rc;
  if (prec == 0 && mma == 0)
    rc = launch_gemm_diag<double, double>(fronts, items, n_items, c0, nb, static_cast<double*>(arena), info, stream);
  else if (prec == 0 && mma == 1)
    rc = launch_gemm_diag<double, float>(fronts, items, n_items, c0, nb, static_cast<double*>(arena), info, stream);
  else if (prec == 1 && mma == 1)
    rc = launch_gemm_diag<float, float>(fronts, items, n_items, c0, nb, static_cast<float*>(arena), info, stream);
  else
    return -1006;
  if (rc) return rc;
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
  else if (prec == 0 && mma == 2)
    rc = launch_gemm_bf16<double>(fronts, items, n_items, c0, nb, schur, transb, tile, static_cast<double*>(arena),
                                  stream);
  else if (prec == 1 && mma == 2)
    rc = launch_gemm_bf16<float>(fronts, items, n_items, c0, nb, schur, transb, tile, static_cast<float*>(arena),
                                 stream);
  else
    return -1006;
  if (rc) return rc;
  return ret(hipGetLastError());
}


long long uhmk_p32_size(int n, int fs, int c0, int nb, int schur) {
  const int kbeg = schur ? 0 : c0, kend = schur ? fs : std::min(c0 + nb, fs);
  if (kend <= kbeg || n <= kend) return 0;
  return p32_ld(n, kend) * (kend - kbeg);
}

int uhmk_p32_pack(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, int schur,
                  const void* arena, int prec, float* P, hipStream_t stream) {
  if (n_items <= 0) return 0;
  if (prec != 0) return -1006;
  hipLaunchKernelGGL(k_p32_pack<double>, dim3(n_items), dim3(256), 0, stream, fronts, items, c0, nb, schur,
                     static_cast<const double*>(arena), P);
  return ret(hipGetLastError());
}

int uhmk_gemm_update_p32(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, int schur,
                         void* arena, int prec, const float* P, hipStream_t stream) {
  if (n_items <= 0) return 0;
  if (prec != 0) return -1006;
  hipLaunchKernelGGL(k_gemm_p32<double>, dim3(n_items), dim3(256), 0, stream, fronts, items, c0, nb, schur,
                     static_cast<double*>(arena), P);
  return ret(hipGetLastError());
}

}  // extern "C"
