// Forward / backward front solves, synthetic element generation and the
// assembled-matrix residual for gfx950 (FP64).
//
// Reference operations restated:
//   forward   op_solve level 1 + solve_lu_piv_1_hier / solve_chol_1_hier
//             UHM/src/operation/element.cxx:243-294,
//             UHM/src/matrix/uhm/fla/lu_piv/solve.cxx:161-187,
//             UHM/src/matrix/uhm/fla/chol/solve.cxx:152-168
//   backward  solve_lu_piv_2_hier / solve_chol_2_hier + Helper_::branch_rhs_x
//             lu_piv/solve.cxx:211-231, chol/solve.cxx:186-202,
//             UHM/include/uhm/matrix/uhm/helper.hxx:302-331
//   rhs merge Helper_::_merge_rhs (helper.hxx:253-301)
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dense_device.hpp"
#include "uhmk.h"

namespace {

// ------------------------------------------------------------------------
// Level-by-level front solves.  Per tree level the forward pass runs
//   gather   x_f = b_f + sum of the children's xb (merge order, through inv);
//            LU: the factor rows are read through the composed pivot
//            permutation, so P is applied while gathering
//   trsv     x_t <- L^-1 x_t on ATL, one workgroup per front
//   gemv     xb -= ABL x_t, 16-row (FP64) / 32-row (FP32) chunks of ABL
// and the backward pass, root first,
//   branch   xb <- the parent's solution rows (emap)
//   gemv     LU: x_t -= ATR xb (row chunks); Cholesky: x_t -= ABL^T xb
//            (column chunks, one wavefront per column)
//   trsv     x_t <- U^-1 x_t (LU) or L^-T x_t (Cholesky)
// Every kernel handles all n_rhs columns (x and b are n x n_rhs per front,
// leading dimension n).  T is the storage type of the level's fronts; all
// arithmetic is FP64.
// ------------------------------------------------------------------------
constexpr int PERM_LDS = 12288;

// perm[i] = row of the unpivoted right-hand side that LAPACK-style swaps
// (k <-> k + p[k], k = 0 .. fs-1) bring to position i.
__global__ __launch_bounds__(64) void k_lu_perm(const uhmk_front* __restrict__ F, const int* __restrict__ flist,
                                                const int* __restrict__ piv, int* __restrict__ perm) {
  __shared__ int s[PERM_LDS];
  const uhmk_front fr = F[flist[blockIdx.x]];
  const int fs = fr.fs;
  const int* __restrict__ p = piv + fr.poff;
  int* __restrict__ out = perm + fr.poff;
  const bool lds = fs <= PERM_LDS;
  int* w = lds ? s : out;
  for (int i = threadIdx.x; i < fs; i += 64) w[i] = i;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < fs; ++k) {
      const int d = p[k];
      if (d) {
        const int t = w[k];
        w[k] = w[k + d];
        w[k + d] = t;
      }
    }
  __syncthreads();
  if (lds)
    for (int i = threadIdx.x; i < fs; i += 64) out[i] = s[i];
}

// item {f, 0, first row, 0}: 256 rows of front f.
template <bool LU>
__global__ __launch_bounds__(256) void k_sv_gather(const uhmk_front* __restrict__ F,
                                                   const uhmk_item* __restrict__ items,
                                                   const int* __restrict__ src_beg,
                                                   const uhmk_source* __restrict__ srcs,
                                                   const int* __restrict__ inv, const int* __restrict__ perm,
                                                   double* __restrict__ x, const double* __restrict__ b, int nrhs) {
  constexpr int MAXS = 8;  // child sources held in registers; more are added in a second pass
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int i = it.b + threadIdx.x;
  if (i >= n) return;
  const int r = (LU && i < fs) ? perm[fr.poff + i] : i;
  double* __restrict__ xo = x + fr.xoff;
  // b is zero on the schur rows of fronts without an input stage (right-hand
  // sides enter at the leaves; refinement places its right-hand side on
  // factor rows only), so those rows skip the read
  const bool rb = fr.soff >= 0 || r < fs;
  const double* __restrict__ bo = b + fr.xoff + r;
  const double* sp[MAXS];
  long long sld[MAXS];
  int ns = 0;
  const int s0 = src_beg[it.f], s1 = src_beg[it.f + 1];
  int s = s0;
  for (; s < s1 && ns < MAXS; ++s) {
    const uhmk_source sc = srcs[s];
    if (sc.src < 0) continue;
    const int ri = inv[sc.inv + r];
    if (ri < 0) continue;
    const uhmk_front ch = F[sc.src];
    sp[ns] = x + ch.xoff + ch.fs + ri;
    sld[ns] = ch.n;
    ++ns;
  }
  // every column written once: b, then the children in merge order
#pragma unroll 4
  for (int c = 0; c < nrhs; ++c) {
    double acc = rb ? bo[(long long)c * n] : 0.0;
#pragma unroll
    for (int q = 0; q < MAXS; ++q)
      if (q < ns) acc += sp[q][(long long)c * sld[q]];
    xo[i + (long long)c * n] = acc;
  }
  for (; s < s1; ++s) {
    const uhmk_source sc = srcs[s];
    if (sc.src < 0) continue;
    const int ri = inv[sc.inv + r];
    if (ri < 0) continue;
    const uhmk_front ch = F[sc.src];
    const double* __restrict__ xc = x + ch.xoff + ch.fs + ri;
    for (int c = 0; c < nrhs; ++c) xo[i + (long long)c * n] += xc[(long long)c * ch.n];
  }
}

// The gather for many right-hand sides: one wavefront per 64-row chunk of a
// front and group of GCW columns, the group's columns held in registers, so
// that every lane has GCW (1 + sources) loads in flight and a small front
// occupies one wavefront, not an idle workgroup.  Workgroup w takes column
// group w % ng of item w / ng: the groups of an item run side by side, so the
// wavefronts in flight sweep a few fronts' contiguous n x n_rhs blocks
// rather than a sliver of every front's block.  Same sums in the same order
// as k_sv_gather: b, then the children in merge order.
constexpr int GCW = 8;

template <bool LU>
__global__ __launch_bounds__(64) void k_sv_gather_cols(const uhmk_front* __restrict__ F,
                                                       const uhmk_item* __restrict__ items,
                                                       const int* __restrict__ src_beg,
                                                       const uhmk_source* __restrict__ srcs,
                                                       const int* __restrict__ inv, const int* __restrict__ perm,
                                                       double* __restrict__ x, const double* __restrict__ b,
                                                       int nrhs, int ng) {
  constexpr int MAXS = 8;
  const uhmk_item it = items[blockIdx.x / ng];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int i = it.b + threadIdx.x;
  if (i >= n) return;
  const int c0 = (blockIdx.x % ng) * GCW;
  const int r = (LU && i < fs) ? perm[fr.poff + i] : i;
  double* __restrict__ xo = x + fr.xoff + i + (long long)c0 * n;
  const bool rb = fr.soff >= 0 || r < fs;
  const double* __restrict__ bo = b + fr.xoff + r + (long long)c0 * n;
  const double* sp[MAXS];
  long long sld[MAXS];
  int ns = 0;
  const int s0 = src_beg[it.f], s1 = src_beg[it.f + 1];
  int s = s0;
  for (; s < s1 && ns < MAXS; ++s) {
    const uhmk_source sc = srcs[s];
    if (sc.src < 0) continue;
    const int ri = inv[sc.inv + r];
    if (ri < 0) continue;
    const uhmk_front ch = F[sc.src];
    sp[ns] = x + ch.xoff + ch.fs + ri + (long long)c0 * ch.n;
    sld[ns] = ch.n;
    ++ns;
  }
  double acc[GCW];
  if (c0 + GCW <= nrhs) {
#pragma unroll
    for (int u = 0; u < GCW; ++u) acc[u] = rb ? bo[(long long)u * n] : 0.0;
#pragma unroll
    for (int q = 0; q < MAXS; ++q)
      if (q < ns) {
#pragma unroll
        for (int u = 0; u < GCW; ++u) acc[u] += sp[q][(long long)u * sld[q]];
      }
  } else {
    const int nc = nrhs - c0;
#pragma unroll
    for (int u = 0; u < GCW; ++u) acc[u] = (rb && u < nc) ? bo[(long long)u * n] : 0.0;
    for (int q = 0; q < ns; ++q)
#pragma unroll
      for (int u = 0; u < GCW; ++u)
        if (u < nc) acc[u] += sp[q][(long long)u * sld[q]];
  }
  for (; s < s1; ++s) {
    const uhmk_source sc = srcs[s];
    if (sc.src < 0) continue;
    const int ri = inv[sc.inv + r];
    if (ri < 0) continue;
    const uhmk_front ch = F[sc.src];
    const double* __restrict__ xc = x + ch.xoff + ch.fs + ri + (long long)c0 * ch.n;
#pragma unroll
    for (int u = 0; u < GCW; ++u)
      if (c0 + u < nrhs) acc[u] += xc[(long long)u * ch.n];
  }
#pragma unroll
  for (int u = 0; u < GCW; ++u)
    if (c0 + u < nrhs) xo[(long long)u * n] = acc[u];
}

// The branch for many right-hand sides, mapped as k_sv_gather_cols: item
// {f, 0, first of 64 rows of xb, 0}.
__global__ __launch_bounds__(64) void k_sv_branch_cols(const uhmk_front* __restrict__ F,
                                                       const uhmk_item* __restrict__ items,
                                                       const int* __restrict__ emap, double* __restrict__ x,
                                                       int nrhs, int ng) {
  const uhmk_item it = items[blockIdx.x / ng];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs, ss = n - fs;
  const int i = it.b + threadIdx.x;
  if (i >= ss || fr.parent < 0) return;
  const int c0 = (blockIdx.x % ng) * GCW;
  const uhmk_front P = F[fr.parent];
  const double* __restrict__ src = x + P.xoff + emap[fr.emap + i] + (long long)c0 * P.n;
  double* __restrict__ dst = x + fr.xoff + fs + i + (long long)c0 * n;
  double v[GCW];
#pragma unroll
  for (int u = 0; u < GCW; ++u) v[u] = (c0 + u < nrhs) ? src[(long long)u * P.n] : 0.0;
#pragma unroll
  for (int u = 0; u < GCW; ++u)
    if (c0 + u < nrhs) dst[(long long)u * n] = v[u];
}

// Broadcast of lane L's double to the wavefront (v_readlane).
template <int L>
__device__ __forceinline__ double bcast(double v) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), L),
                          __builtin_amdgcn_readlane(__double2loint(v), L));
}

// One wavefront solves a 64 x 64 triangle held one entry vector per lane:
// step K finalises x_K (times rd when SCALE) in its owner lane, broadcasts it
// and updates x_j -= e[K] x_K in the other lanes, where e holds the lane's
// strict-triangle entries (zero elsewhere).  ASC runs K = 0 .. 63, else
// 63 .. 0; steps with K >= kb are skipped.
template <int K, bool ASC, bool SCALE>
struct TrsvStep {
  __device__ __forceinline__ static void run(double& xj, const double (&e)[64], double rd, int lane, int kb) {
    constexpr int KK = ASC ? K : 63 - K;
    if (KK < kb) {
      if (SCALE && lane == KK) xj *= rd;
      const double xk = bcast<KK>(xj);
      xj = fma(-e[KK], xk, xj);
    }
    TrsvStep<K + 1, ASC, SCALE>::run(xj, e, rd, lane, kb);
  }
};
template <bool ASC, bool SCALE>
struct TrsvStep<64, ASC, SCALE> {
  __device__ __forceinline__ static void run(double&, const double (&)[64], double, int, int) {}
};

template <typename T>
__device__ __forceinline__ double ld(const T* p) {
  return (double)*p;
}

// Triangular solve of the diagonal super-block [s0, s1) of ATL, s0 = sb * SB,
// s1 = min(s0 + SB, fs), SB <= 1024; fronts with fs <= s0 are skipped.
// Forward: x <- L^-1 x (LU: unit lower; Cholesky: lower).  Backward:
// x <- U^-1 x (LU) or L^-T x (Cholesky).  The couplings with the rest of
// ATL are applied by the gemv kernels between super-blocks.  One workgroup
// of NT threads per front; the super-block's part of x lives in LDS and is
// processed in 64-row blocks.  With INV (Cholesky factors from 64-wide
// panels) the diagonal block holds W = L(k,k)^-1 in its strict upper
// triangle (transposed) and L's diagonal, and the block step is a GEMV with
// W (forward) or W^T (backward); otherwise one wavefront solves the 64 x 64
// triangle step by step (TrsvStep).  Each block's matrix loads are issued
// before its barriers so that they overlap.
constexpr int SB_MAX = 1024;

template <typename T, bool CHOL, bool FWD, bool INV, int NT>
__global__ __launch_bounds__(NT) void k_sv_trsv(const uhmk_front* __restrict__ F, const int* __restrict__ flist,
                                                const T* __restrict__ arena, double* __restrict__ x, int nrhs,
                                                int sb, int SB) {
  constexpr bool STAGE = INV || (!FWD && CHOL);
  constexpr int NPRE = 64 * 64 / NT;
  __shared__ double sx[SB_MAX];
  __shared__ double sL[STAGE ? 64 * 65 : 1];
  const uhmk_front fr = F[flist[blockIdx.x]];
  const int n = fr.n, fs = fr.fs;
  const int s0 = sb * SB;
  if (s0 >= fs) return;
  const int s1 = min(s0 + SB, fs), m = s1 - s0;
  const T* __restrict__ A = arena + fr.off;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  constexpr int NW = NT / 64;
  const int nblk = (m + 63) / 64;
  for (int c = 0; c < nrhs; ++c) {
    double* __restrict__ xg = x + fr.xoff + (long long)c * n + s0;
    for (int i = tid; i < m; i += NT) sx[i] = xg[i];
    __syncthreads();
    for (int bq = 0; bq < nblk; ++bq) {
      const int bk = FWD ? bq : nblk - 1 - bq;
      const int l0 = 64 * bk, kb = min(64, m - l0), l1 = l0 + kb;  // local rows
      const int k0 = s0 + l0;                                       // absolute
      // ---- loads that do not depend on x
      double pre[STAGE ? NPRE : 1];
      if (STAGE) {
#pragma unroll
        for (int u = 0; u < NPRE; ++u) {
          const int idx = tid + NT * u, r = idx & 63, cc = idx >> 6;
          pre[u] = uhmd::ldz(A, (k0 + r) + (long long)(k0 + cc) * n, r < kb && cc < kb);
        }
      }
      if (!FWD && CHOL) {
        // x_q -= sum_{i in [l1, m)} L(i, q) x_i, NW columns at a time per wave
        double acc[64 / NW];
#pragma unroll
        for (int u = 0; u < 64 / NW; ++u) acc[u] = 0.0;
#pragma unroll 4
        for (int i = l1 + lane; i < m; i += 64) {
          const double xi = sx[i];
#pragma unroll
          for (int u = 0; u < 64 / NW; ++u) {
            const int q = wid + NW * u;
            if (q < kb) acc[u] = fma(ld(A + (s0 + i) + (long long)(k0 + q) * n), xi, acc[u]);
          }
        }
#pragma unroll
        for (int u = 0; u < 64 / NW; ++u) {
          double a = acc[u];
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
          const int q = wid + NW * u;
          if (lane == 0 && q < kb) sx[l0 + q] -= a;
        }
      }
      if (STAGE) {
#pragma unroll
        for (int u = 0; u < NPRE; ++u) {
          const int idx = tid + NT * u, r = idx & 63, cc = idx >> 6;
          sL[r * 65 + cc] = pre[u];
        }
        __syncthreads();
      }
      if (tid < 64) {
        const int j = lane;
        if (INV) {
          // block entry (k, j), k < j: W(j, k); diagonal: L(j, j)
          double y = 0.0;
          if (j < kb) {
            if (FWD) {
              y = sx[l0 + j] / sL[j * 65 + j];
              for (int k = 0; k < j; ++k) y = fma(sL[k * 65 + j], sx[l0 + k], y);
            } else {
              y = sx[l0 + j] / sL[j * 65 + j];
              for (int k = j + 1; k < kb; ++k) y = fma(sL[j * 65 + k], sx[l0 + k], y);
            }
          }
          __builtin_amdgcn_wave_barrier();
          if (j < kb) sx[l0 + j] = y;
        } else {
          double e[64];
          double rd = 1.0;
          if (FWD) {
            // row j of the block, strict lower part
#pragma unroll
            for (int q = 0; q < 64; ++q) e[q] = uhmd::ldz(A, (k0 + j) + (long long)(k0 + q) * n, j < kb && q < j);
            if (CHOL && j < kb) rd = 1.0 / ld(A + (k0 + j) + (long long)(k0 + j) * n);
          } else if (CHOL) {
            // column j of the block below the diagonal: e[i] = L(i, j), i > j
#pragma unroll
            for (int q = 0; q < 64; ++q) e[q] = (q > j) ? sL[q * 65 + j] : 0.0;
            if (j < kb) rd = 1.0 / sL[j * 65 + j];
          } else {
            // row j of U, strict upper part
#pragma unroll
            for (int q = 0; q < 64; ++q)
              e[q] = uhmd::ldz(A, (k0 + j) + (long long)(k0 + q) * n, j < kb && q > j && q < kb);
            if (j < kb) rd = 1.0 / ld(A + (k0 + j) + (long long)(k0 + j) * n);
          }
          double xj = (j < kb) ? sx[l0 + j] : 0.0;
          if (FWD) {
            if (CHOL) TrsvStep<0, true, true>::run(xj, e, rd, lane, kb);
            else TrsvStep<0, true, false>::run(xj, e, rd, lane, kb);
          } else {
            TrsvStep<0, false, true>::run(xj, e, rd, lane, kb);
          }
          if (j < kb) sx[l0 + j] = xj;
        }
      }
      __syncthreads();
      if (FWD || !CHOL) {
        // FWD: rows below the block; LU backward: rows above it
        const int r0 = FWD ? l1 : 0, r1 = FWD ? m : l0;
        for (int i = r0 + tid; i < r1; i += NT) {
          const T* __restrict__ Ai = A + (s0 + i) + (long long)k0 * n;
          double acc = 0.0;
          if (kb == 64) {
            double e[64];
#pragma unroll
            for (int q = 0; q < 64; ++q) e[q] = ld(Ai + (long long)q * n);
#pragma unroll
            for (int q = 0; q < 64; ++q) acc = fma(e[q], sx[l0 + q], acc);
          } else {
            for (int q = 0; q < kb; ++q) acc = fma(ld(Ai + (long long)q * n), sx[l0 + q], acc);
          }
          sx[i] -= acc;
        }
        __syncthreads();
      }
    }
    for (int i = tid; i < m; i += NT) xg[i] = sx[i];
    __syncthreads();
  }
}

// Row-chunk GEMV: x[rows] -= A[rows, cbeg:cend) x[cbeg:cend) over RT rows
// per item (one 128-byte line of each column), rows [r0, min(r0 + RT,
// r0 < fs ? fs : n)).  item {f, cbeg, r0, cend}.  Serves xb -= ABL x_t,
// x_t -= ATR xb and the super-block couplings inside ATL.
template <typename T>
__global__ __launch_bounds__(256) void k_sv_gemv_rows(const uhmk_front* __restrict__ F,
                                                      const uhmk_item* __restrict__ items,
                                                      const T* __restrict__ arena, double* __restrict__ x, int nrhs) {
  constexpr int RT = 128 / (int)sizeof(T);
  constexpr int KG = 64 / RT;
  __shared__ double red[4][RT];
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const T* __restrict__ A = arena + fr.off;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r = lane % RT, kg = lane / RT;
  const int rend = min(it.b + RT, it.b < fs ? fs : n);
  const int cbeg = it.a, cend = it.c;
  const int i = it.b + r;
  const bool live = i < rend;
  for (int c = 0; c < nrhs; ++c) {
    double* __restrict__ xv = x + fr.xoff + (long long)c * n;
    double acc = 0.0;
    if (live) {
#pragma unroll 16
      for (int k = cbeg + w * KG + kg; k < cend; k += 4 * KG) acc = fma(ld(A + i + (long long)k * n), xv[k], acc);
    }
#pragma unroll
    for (int o = RT; o < 64; o <<= 1) acc += __shfl_xor(acc, o);
    if (kg == 0) red[w][r] = acc;
    __syncthreads();
    if (tid < RT && it.b + tid < rend) xv[it.b + tid] -= (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
    __syncthreads();
  }
}

// Column-chunk transposed GEMV: x[col] -= sum_{i in [rbeg, rend)} A(i, col)
// x[i] for 16 columns [c0, min(c0 + 16, fs)), four per wavefront at once.
// item {f, rbeg, c0, rend}; the result goes to dst (the same front layout
// as x; dst may be x).  Serves x_t -= ABL^T xb and the super-block
// couplings of the Cholesky backward solve.
template <typename T>
__global__ __launch_bounds__(256) void k_sv_gemv_cols(const uhmk_front* __restrict__ F,
                                                      const uhmk_item* __restrict__ items,
                                                      const T* __restrict__ arena, double* __restrict__ x, int nrhs,
                                                      double* __restrict__ dst) {
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const T* __restrict__ A = arena + fr.off;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int cend = min(it.b + 16, fs);
  for (int c = 0; c < nrhs; ++c) {
    double* __restrict__ xv = x + fr.xoff + (long long)c * n;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int i = it.a + lane; i < it.c; i += 64) {
      const double xi = xv[i];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int col = it.b + w + 4 * u;
        if (col < cend) acc[u] = fma(ld(A + i + (long long)col * n), xi, acc[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      double a = acc[u];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
      const int col = it.b + w + 4 * u;
      if (lane == 0 && col < cend) dst[fr.xoff + (long long)c * n + col] -= a;
    }
  }
}

// ------------------------------------------------------------------------
// Cholesky solves by 64-row block steps, for factors whose 64 x 64 diagonal
// blocks carry W = L(k,k)^-1 (strict upper triangle, transposed; diagonal
// L(k,k)).  One launch per block step k0 of a tree level; every workgroup
// forms y = W x_k (forward) or y = W^T z_k (backward) itself and then
// updates one 64-row (forward) or 64-column (backward) chunk, so a big
// front is spread over many workgroups.  Two front-layout vectors keep the
// steps free of read/write overlap: forward reads x, writes the solved
// block to z and updates x below it; backward reads z, writes the solved
// block to x and updates z above it.
// ------------------------------------------------------------------------
// The kb x kb diagonal block at D (column stride n) into registers (thread
// tid: entries tid + 256 u, row-fastest), then into LDS as sW[r * 65 + c];
// split so that the loads can be issued long before the block is needed.
template <typename T>
__device__ __forceinline__ void load_diag_block(double (&v)[16], const T* __restrict__ D, long long n, int kb,
                                                int tid) {
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int idx = tid + 256 * u, r = idx & 63, c = idx >> 6;
    v[u] = uhmd::ldz(D, r + (long long)c * n, r < kb && c < kb);
  }
}
__device__ __forceinline__ void store_diag_block(double* __restrict__ sW, const double (&v)[16], int tid) {
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int idx = tid + 256 * u, r = idx & 63, c = idx >> 6;
    sW[r * 65 + c] = v[u];
  }
}
template <typename T>
__device__ __forceinline__ void stage_diag_block(double* __restrict__ sW, const T* __restrict__ D, long long n, int kb,
                                                 int tid) {
  double v[16];
  load_diag_block(v, D, n, kb, tid);
  store_diag_block(sW, v, tid);
}

// item {f, 0, first row of the chunk, 0}: r0 == k0 is the block itself,
// the others start at k0 + kb, k0 + kb + 64, ... and run into ABL.
template <typename T>
__global__ __launch_bounds__(256) void k_sv_chol_fwd_step(const uhmk_front* __restrict__ F,
                                                          const uhmk_item* __restrict__ items,
                                                          const T* __restrict__ arena, double* __restrict__ x,
                                                          double* __restrict__ z, int nrhs, int k0) {
  __shared__ double sW[64 * 65];
  __shared__ double sy[64];
  __shared__ double sp[4][64];
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int kb = min(64, fs - k0);
  const T* __restrict__ A = arena + fr.off;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r0 = it.b;
  const bool own = (r0 == k0);
  // chunk rows: this thread's row lane, quarter w of the 64 columns
  const int row = r0 + lane;
  double e[16];
  if (!own) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = 16 * w + q;
      e[q] = uhmd::ldz(A, row + (long long)(k0 + k) * n, row < n && k < kb);
    }
  }
  stage_diag_block(sW, A + k0 + (long long)k0 * n, n, kb, tid);
  __syncthreads();
  __shared__ double sv[64];
  for (int c = 0; c < nrhs; ++c) {
    if (tid < 64) sv[tid] = uhmd::ldz(x + fr.xoff + (long long)c * n + k0, tid, tid < kb);
    __syncthreads();
    const double* __restrict__ xs = sv;
    // y_j = x_j / L(j,j) + sum_{k<j} W(j,k) x_k; four threads per j
    {
      const int j = lane, part = w;
      double acc = 0.0;
      if (j < kb) {
        for (int k = 16 * part; k < 16 * part + 16 && k < j; ++k) acc = fma(sW[k * 65 + j], xs[k], acc);
        if (part == 0) acc += xs[j] / sW[j * 65 + j];
      }
      sp[part][j] = acc;
    }
    __syncthreads();
    if (tid < 64) sy[tid] = (sp[0][tid] + sp[1][tid]) + (sp[2][tid] + sp[3][tid]);
    __syncthreads();
    if (own) {
      if (tid < kb) z[fr.xoff + (long long)c * n + k0 + tid] = sy[tid];
    } else {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < 16; ++q) acc = fma(e[q], sy[16 * w + q], acc);
      sp[w][lane] = acc;
      __syncthreads();
      if (tid < 64 && row < n)
        x[fr.xoff + (long long)c * n + row] -= (sp[0][tid] + sp[1][tid]) + (sp[2][tid] + sp[3][tid]);
    }
    __syncthreads();
  }
}

// item {f, 0, first column of the chunk (<= k0), 0}; the chunk starting at
// k0 is the block itself, the others lie left of it.
template <typename T>
__global__ __launch_bounds__(256) void k_sv_chol_bwd_step(const uhmk_front* __restrict__ F,
                                                          const uhmk_item* __restrict__ items,
                                                          const T* __restrict__ arena, double* __restrict__ x,
                                                          double* __restrict__ z, int nrhs, int k0) {
  __shared__ double sW[64 * 65];
  __shared__ double sT[64 * 65];
  __shared__ double sy[64];
  __shared__ double sp[4][64];
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int kb = min(64, fs - k0);
  const T* __restrict__ A = arena + fr.off;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int c0 = it.b;
  const bool own = (c0 == k0);
  if (!own) {
    // L(k0 + r, c0 + c): rows of the block, columns of the chunk
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int idx = tid + 256 * u, r = idx & 63, cc = idx >> 6;
      v[u] = uhmd::ldz(A, (k0 + r) + (long long)(c0 + cc) * n, r < kb);
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int idx = tid + 256 * u, r = idx & 63, cc = idx >> 6;
      sT[r * 65 + cc] = v[u];
    }
  }
  stage_diag_block(sW, A + k0 + (long long)k0 * n, n, kb, tid);
  __syncthreads();
  __shared__ double sv[64];
  for (int c = 0; c < nrhs; ++c) {
    if (tid < 64) sv[tid] = uhmd::ldz(z + fr.xoff + (long long)c * n + k0, tid, tid < kb);
    __syncthreads();
    const double* __restrict__ zs = sv;
    // y_j = z_j / L(j,j) + sum_{k>j} W(k,j) z_k; W(k, j) is block entry (j, k)
    {
      const int j = lane, part = w;
      double acc = 0.0;
      if (j < kb) {
        for (int k = max(j + 1, 16 * part); k < 16 * part + 16 && k < kb; ++k) acc = fma(sW[j * 65 + k], zs[k], acc);
        if (part == 0) acc += zs[j] / sW[j * 65 + j];
      }
      sp[part][j] = acc;
    }
    __syncthreads();
    if (tid < 64) sy[tid] = (sp[0][tid] + sp[1][tid]) + (sp[2][tid] + sp[3][tid]);
    __syncthreads();
    if (own) {
      if (tid < kb) x[fr.xoff + (long long)c * n + k0 + tid] = sy[tid];
    } else {
      // z(c0 + cc) -= sum_r L(k0 + r, c0 + cc) y_r; thread (w, lane): column
      // lane, rows 16 w .. 16 w + 15
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < 16; ++q) acc = fma(sT[(16 * w + q) * 65 + lane], sy[16 * w + q], acc);
      sp[w][lane] = acc;
      __syncthreads();
      if (tid < 64) z[fr.xoff + (long long)c * n + c0 + tid] -= (sp[0][tid] + sp[1][tid]) + (sp[2][tid] + sp[3][tid]);
    }
    __syncthreads();
  }
}

// Branch: xb <- the parent's solution at emap.  item {f, 0, first row, 0}.
__global__ __launch_bounds__(256) void k_sv_branch(const uhmk_front* __restrict__ F,
                                                   const uhmk_item* __restrict__ items,
                                                   const int* __restrict__ emap, double* __restrict__ x, int nrhs) {
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs, ss = n - fs;
  const int i = it.b + threadIdx.x;
  if (i >= ss || fr.parent < 0) return;
  const uhmk_front P = F[fr.parent];
  const int pi = emap[fr.emap + i];
  const double* src = x + P.xoff + pi;
  double* dst = x + fr.xoff + fs + i;
  int c = 0;
  // four columns' loads in flight before their stores (parent and child
  // rows never overlap)
  for (; c + 4 <= nrhs; c += 4) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = src[(long long)(c + u) * P.n];
#pragma unroll
    for (int u = 0; u < 4; ++u) dst[(long long)(c + u) * n] = v[u];
  }
  for (; c < nrhs; ++c) dst[(long long)c * n] = src[(long long)c * P.n];
}

// ------------------------------------------------------------------------
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) counter-based RNG and
// Box-Muller normals: identical streams for every launch order.
// ------------------------------------------------------------------------
struct U4 {
  unsigned x, y, z, w;
};

__device__ inline U4 philox(U4 c, unsigned k0, unsigned k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const unsigned long long p0 = (unsigned long long)0xD2511F53u * c.x;
    const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c.z;
    const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
    const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

__device__ inline double rng_normal(unsigned long long seed, unsigned elem, unsigned stream, unsigned idx) {
  const U4 r = philox(U4{idx >> 1, elem, stream, 0u}, (unsigned)seed, (unsigned)(seed >> 32));
  const double u1 = ((double)r.x + 1.0) * 2.3283064365386963e-10;  // (0, 1]
  const double u2 = (double)r.y * 2.3283064365386963e-10;          // [0, 1)
  const double rad = sqrt(-2.0 * log(u1));
  const double th = 6.283185307179586 * u2;
  return (idx & 1) ? rad * sin(th) : rad * cos(th);
}

__device__ inline double rng_uniform(unsigned long long seed, unsigned elem, unsigned stream, unsigned idx) {
  const U4 r = philox(U4{idx, elem, stream, 1u}, (unsigned)seed, (unsigned)(seed >> 32));
  return ((double)r.x + 0.5) * 2.3283064365386963e-10;
}

// Synthetic element matrices K_e = s_e * (M^T M + alpha I + beta S) with
// M, S i.i.d. N(0,1) (m x m) and the per-element scale s_e drawn by kind:
// 0 -> 1, 1 -> exp(p0 * N(0,1)) (lognormal(0, p0)), 2 -> 10^(p0 + (p1-p0) U).
// Written straight into the leaf stage in front order through lmap, and
// b_e ~ N(0,1) into the right-hand side.  Item: f, a = element id,
// b = lmap offset, c = m.
__global__ __launch_bounds__(256) void k_gen_elements(const uhmk_front* __restrict__ F,
                                                      const uhmk_item* __restrict__ items,
                                                      const int* __restrict__ lmap,
                                                      unsigned long long seed, double alpha,
                                                      double beta, int scale_kind, double p0,
                                                      double p1, double* __restrict__ stage,
                                                      double* __restrict__ b, int nrhs, int with_lhs,
                                                      int with_rhs) {
  extern __shared__ double sM[];
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int m = it.c, ldm = m | 1, n = fr.n;
  const unsigned eid = (unsigned)it.a;
  const int* __restrict__ mp = lmap + it.b;
  const int tid = threadIdx.x;
  if (with_lhs) {
    for (int idx = tid; idx < m * m; idx += 256) {
      const int i = idx % m, j = idx / m;
      sM[i + j * ldm] = rng_normal(seed, eid, 0u, (unsigned)idx);
    }
    __syncthreads();
    double s = 1.0;
    if (scale_kind == 1) s = exp(p0 * rng_normal(seed, eid, 2u, 0u));
    else if (scale_kind == 2) s = pow(10.0, p0 + (p1 - p0) * rng_uniform(seed, eid, 2u, 0u));
    // alpha < 0: the shift is |alpha| times the element size m
    const double shift = alpha < 0.0 ? -alpha * (double)m : alpha;
    for (int idx = tid; idx < m * m; idx += 256) {
      const int i = idx % m, j = idx / m;
      double acc = 0.0;
      for (int k = 0; k < m; ++k) acc += sM[k + i * ldm] * sM[k + j * ldm];
      if (i == j) acc += shift;
      if (beta != 0.0) acc += beta * rng_normal(seed, eid, 1u, (unsigned)idx);
      stage[fr.soff + mp[i] + (long long)mp[j] * n] = s * acc;
    }
  }
  if (with_rhs) {
    for (int idx = tid; idx < m * nrhs; idx += 256) {
      const int i = idx % m, r = idx / m;
      b[fr.xoff + mp[i] + (long long)r * n] = rng_normal(seed, eid, 3u, (unsigned)idx);
    }
  }
}

// Assembled values: val[p] = sum of the leaf-stage entries listed for slot
// p, in leaf order (deterministic).
__global__ __launch_bounds__(256) void k_csr_values(long long nnz, const long long* __restrict__ cptr,
                                                    const long long* __restrict__ cidx,
                                                    const double* __restrict__ stage, double* __restrict__ val) {
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  if (p >= nnz) return;
  double acc = 0.0;
  for (long long q = cptr[p]; q < cptr[p + 1]; ++q) acc += stage[cidx[q]];
  val[p] = acc;
}

// Assembled right-hand side column: bg[g] = sum of the leaf rhs entries of
// global row g (b[bidx[q] + col * bn[q]]).
__global__ __launch_bounds__(256) void k_gather_rhs(long long nrow, const long long* __restrict__ bptr,
                                                    const long long* __restrict__ bidx, const int* __restrict__ bn,
                                                    const double* __restrict__ b, double* __restrict__ bg, int col) {
  const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
  if (g >= nrow) return;
  double acc = 0.0;
  for (long long q = bptr[g]; q < bptr[g + 1]; ++q) acc += b[bidx[q] + (long long)col * bn[q]];
  bg[g] = acc;
}

__global__ __launch_bounds__(256) void k_fill_int(int* __restrict__ p, long long n, int v) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------
// Cholesky level solves in one launch per tree level and direction, for
// factors whose 64 x 64 diagonal blocks carry W = L_kk^-1 (the solves of
// solve_chol_{1,2}_hier, UHM/src/matrix/uhm/fla/chol/solve.cxx:152-202,
// for every front of a level at once).  One workgroup per 64-row block of
// ATL, left-looking: forward, block r applies L(r, j) y_j of every earlier
// block j as soon as y_j appears, then y_r = W_r (x_r - sum) goes to z;
// workgroups for 64-row chunks of ABL apply the same terms to xb.
// Backward, block r (listed in decreasing order) forms
// z_r - ABL(:, r)^T xb - sum_{j > r} L(j, r)^T x_j and x_r = W_r^T of it.
// A block's result is read by the others of the launch while they run:
// the solution blocks start as SOLVE_MARK (a NaN bit pattern that no
// arithmetic produces: z by uhmk_solve_mark before the forward pass, the
// ATL rows of x by the forward pass itself once it has read them), the
// result is stored with agent-scope coherent stores and the readers poll
// their values with coherent loads until the mark is gone (one memory round
// trip per dependency, no flags).  Polls are bounded: after SOLVE_SPIN
// rounds *err is raised and the launch finishes with invalid values
// instead of hanging.  Every wait is on a block of the same front listed
// earlier in the launch, so the in-order dispatch of workgroups leaves the
// awaited block running or done.
// ------------------------------------------------------------------------
constexpr unsigned long long SOLVE_MARK = 0xFFF4B17E5E4D1A00ULL;
constexpr int SOLVE_SPIN = 1 << 22;
constexpr int SOLVE_MAXR = 8;  // right-hand sides per level launch

__device__ __forceinline__ bool is_mark(double v) {
  return (unsigned long long)__double_as_longlong(v) == SOLVE_MARK;
}

__device__ __forceinline__ double poll_value(double* p, int* err) {
  double v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int spins = 0;
  while (is_mark(v)) {
    if ((++spins & 1023) == 0 &&
        (spins > SOLVE_SPIN || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      if (spins > SOLVE_SPIN) atomicAdd(err, 1);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
    v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return v;
}

__device__ __forceinline__ void store_coh(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// item {f, 0 ATL block / 1 ABL chunk, first row, 0}; thread (w, lane): row
// r0 + lane, columns 16 w .. 16 w + 15 of each block term
template <typename T>
__global__ __launch_bounds__(256) void k_sv_chol_fwd_level(const uhmk_front* __restrict__ F,
                                                           const uhmk_item* __restrict__ items,
                                                           const T* __restrict__ arena, double* __restrict__ x,
                                                           double* __restrict__ z, int nrhs, int* __restrict__ err) {
  __shared__ double sW[64 * 65];
  __shared__ double sy[SOLVE_MAXR][64];
  __shared__ double sp[SOLVE_MAXR][4][64];
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs, r0 = it.b;
  const bool atl = it.a == 0;
  const T* __restrict__ A = arena + fr.off;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int row = r0 + lane;
  const int rend = atl ? min(r0 + 64, fs) : n;
  const int nterm = atl ? r0 / 64 : (fs + 63) / 64;
  double* __restrict__ xf = x + fr.xoff;
  double* __restrict__ zf = z + fr.xoff;
  double part[SOLVE_MAXR];
#pragma unroll
  for (int c = 0; c < SOLVE_MAXR; ++c) part[c] = 0.0;
  // the diagonal block (W) of an ATL block goes to LDS before the terms
  // (sW is not used by them), so that its load is off the chain from the
  // previous block's result
  if (atl) stage_diag_block(sW, A + r0 + (long long)r0 * n, n, rend - r0, tid);
  // the terms' L(r, j) entries are loaded one term ahead
  double e[16], en[16];
  auto load_term = [&](double (&d)[16], int j) {
    const int c0 = 64 * j, kbj = min(64, fs - c0);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = 16 * w + q;
      d[q] = uhmd::ldz(A, row + (long long)(c0 + k) * n, row < rend && k < kbj);
    }
  };
  if (nterm > 0) load_term(e, 0);
  for (int j = 0; j < nterm; ++j) {
    const int c0 = 64 * j, kbj = min(64, fs - c0);
    if (j > 0) {
#pragma unroll
      for (int q = 0; q < 16; ++q) e[q] = en[q];
    }
    if (j + 1 < nterm) load_term(en, j + 1);
    if (tid < 64)
      for (int c = 0; c < nrhs; ++c) sy[c][tid] = tid < kbj ? poll_value(zf + (long long)c * n + c0 + tid, err) : 0.0;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < SOLVE_MAXR; ++c)
      if (c < nrhs) {
#pragma unroll
        for (int q = 0; q < 16; ++q) part[c] = fma(e[q], sy[c][16 * w + q], part[c]);
      }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < SOLVE_MAXR; ++c)
    if (c < nrhs) sp[c][w][lane] = part[c];
  __syncthreads();
  if (!atl) {
    if (tid < 64 && row < n)
      for (int c = 0; c < nrhs; ++c)
        xf[(long long)c * n + row] -= (sp[c][0][tid] + sp[c][1][tid]) + (sp[c][2][tid] + sp[c][3][tid]);
    return;
  }
  const int kb = rend - r0;
  if (tid < 64)
    for (int c = 0; c < nrhs; ++c)
      sy[c][tid] = tid < kb ? xf[(long long)c * n + row] -
                                  ((sp[c][0][tid] + sp[c][1][tid]) + (sp[c][2][tid] + sp[c][3][tid]))
                            : 0.0;
  __syncthreads();
  // y_j = acc_j / L(j,j) + sum_{k<j} W(j,k) acc_k, four threads per j
  for (int c = 0; c < nrhs; ++c) {
    const int jj = lane, q4 = w;
    double acc = 0.0;
    if (jj < kb) {
      for (int k = 16 * q4; k < 16 * q4 + 16 && k < jj; ++k) acc = fma(sW[k * 65 + jj], sy[c][k], acc);
      if (q4 == 0) acc += sy[c][jj] / sW[jj * 65 + jj];
    }
    sp[c][q4][jj] = acc;
  }
  __syncthreads();
  if (tid < kb)
    for (int c = 0; c < nrhs; ++c) {
      store_coh(zf + (long long)c * n + row,
                (sp[c][0][tid] + sp[c][1][tid]) + (sp[c][2][tid] + sp[c][3][tid]));
      // the block's rows of x are read: they become the backward pass's mark
      xf[(long long)c * n + row] = __longlong_as_double((long long)SOLVE_MARK);
    }
}

// item {f, 0, first row of the ATL block, 0}, each front's blocks in
// decreasing order.  Every 64 x 64 piece of the block's columns (ABL rows,
// then the ATL blocks below) is staged through LDS with coalesced loads
// (thread t: row t % 64, columns t / 64 + 4 u), the next one loaded while
// the current one is applied; thread (w, lane) then sums column lane over
// rows 16 w .. 16 w + 15 of the piece.
template <typename T>
__global__ __launch_bounds__(256) void k_sv_chol_bwd_level(const uhmk_front* __restrict__ F,
                                                           const uhmk_item* __restrict__ items,
                                                           const T* __restrict__ arena, double* __restrict__ x,
                                                           double* __restrict__ z, int nrhs, int* __restrict__ err) {
  __shared__ double sW[64 * 65];  // the piece, then the diagonal block
  __shared__ double sy[SOLVE_MAXR][64];
  __shared__ double sp[SOLVE_MAXR][4][64];
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs, r0 = it.b;
  const int kb = min(64, fs - r0);
  const T* __restrict__ A = arena + fr.off;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int col = r0 + lane;
  double* __restrict__ xf = x + fr.xoff;
  double* __restrict__ zf = z + fr.xoff;
  double part[SOLVE_MAXR];
#pragma unroll
  for (int c = 0; c < SOLVE_MAXR; ++c) part[c] = 0.0;
  // pieces: ABL rows [fs, n) in 64-row chunks, then the ATL blocks below
  // this one from the last (the first to be solved) upwards
  const int nabl = (n - fs + 63) / 64, nblk = (fs + 63) / 64, rb = r0 / 64;
  const int npiece = nabl + (nblk - 1 - rb);
  auto piece_row = [&](int pc) { return pc < nabl ? fs + 64 * pc : 64 * (nblk - 1 - (pc - nabl)); };
  double dv[16];
  load_diag_block(dv, A + r0 + (long long)r0 * n, n, kb, tid);
  double v[16];
  auto load_piece = [&](int pc) {
    const int p0 = piece_row(pc), pend = pc < nabl ? n : min(p0 + 64, fs);
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int cc = w + 4 * u;
      v[u] = uhmd::ldz(A, (p0 + lane) + (long long)(r0 + cc) * n, p0 + lane < pend && cc < kb);
    }
  };
  if (npiece > 0) load_piece(0);
  for (int pc = 0; pc < npiece; ++pc) {
    const int p0 = piece_row(pc);
    const bool abl = pc < nabl;
    const int pend = abl ? n : min(p0 + 64, fs);
    // the piece into LDS (sW[row][column]), the next piece's loads issued
#pragma unroll
    for (int u = 0; u < 16; ++u) sW[lane * 65 + w + 4 * u] = v[u];
    if (pc + 1 < npiece) load_piece(pc + 1);
    // the multiplying values: xb rows of the chunk, or the solved block
    if (tid < 64)
      for (int c = 0; c < nrhs; ++c)
        sy[c][tid] = p0 + tid >= pend ? 0.0
                     : abl         ? xf[(long long)c * n + p0 + tid]
                                   : poll_value(xf + (long long)c * n + p0 + tid, err);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < SOLVE_MAXR; ++c)
      if (c < nrhs) {
#pragma unroll
        for (int q = 0; q < 16; ++q) part[c] = fma(sW[(16 * w + q) * 65 + lane], sy[c][16 * w + q], part[c]);
      }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < SOLVE_MAXR; ++c)
    if (c < nrhs) sp[c][w][lane] = part[c];
  store_diag_block(sW, dv, tid);
  __syncthreads();
  if (tid < 64)
    for (int c = 0; c < nrhs; ++c)
      sy[c][tid] = tid < kb ? zf[(long long)c * n + col] -
                                  ((sp[c][0][tid] + sp[c][1][tid]) + (sp[c][2][tid] + sp[c][3][tid]))
                            : 0.0;
  __syncthreads();
  // x_j = acc_j / L(j,j) + sum_{k>j} W(k,j) acc_k; W(k,j) is block entry (j, k)
  for (int c = 0; c < nrhs; ++c) {
    const int jj = lane, q4 = w;
    double acc = 0.0;
    if (jj < kb) {
      for (int k = max(jj + 1, 16 * q4); k < 16 * q4 + 16 && k < kb; ++k) acc = fma(sW[jj * 65 + k], sy[c][k], acc);
      if (q4 == 0) acc += sy[c][jj] / sW[jj * 65 + jj];
    }
    sp[c][q4][jj] = acc;
  }
  __syncthreads();
  if (tid < kb)
    for (int c = 0; c < nrhs; ++c)
      store_coh(xf + (long long)c * n + col, (sp[c][0][tid] + sp[c][1][tid]) + (sp[c][2][tid] + sp[c][3][tid]));
}

__global__ __launch_bounds__(256) void k_fill_mark(double* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = __longlong_as_double((long long)SOLVE_MARK);
}

inline int ret(hipError_t e) { return e == hipSuccess ? 0 : -(int)e; }

template <typename T, bool CHOL, bool FWD, bool INV>
int launch_trsv(const uhmk_front* fronts, const int* flist, int n_fronts, int max_fs, const T* arena, double* x,
                int nrhs, int sb, int SB, hipStream_t stream) {
  if (SB > SB_MAX || SB < 64 || SB % 64) return -1008;
  if (std::min(max_fs, SB) > 256)
    hipLaunchKernelGGL((k_sv_trsv<T, CHOL, FWD, INV, 512>), dim3(n_fronts), dim3(512), 0, stream, fronts, flist, arena,
                       x, nrhs, sb, SB);
  else
    hipLaunchKernelGGL((k_sv_trsv<T, CHOL, FWD, INV, 256>), dim3(n_fronts), dim3(256), 0, stream, fronts, flist, arena,
                       x, nrhs, sb, SB);
  return ret(hipGetLastError());
}

template <typename T>
int trsv_any(int chol, int fwd, int inv, const uhmk_front* fronts, const int* flist, int n_fronts, int max_fs,
             const T* arena, double* x, int nrhs, int sb, int SB, hipStream_t stream) {
  if (chol && inv)
    return fwd ? launch_trsv<T, true, true, true>(fronts, flist, n_fronts, max_fs, arena, x, nrhs, sb, SB, stream)
               : launch_trsv<T, true, false, true>(fronts, flist, n_fronts, max_fs, arena, x, nrhs, sb, SB, stream);
  if (chol)
    return fwd ? launch_trsv<T, true, true, false>(fronts, flist, n_fronts, max_fs, arena, x, nrhs, sb, SB, stream)
               : launch_trsv<T, true, false, false>(fronts, flist, n_fronts, max_fs, arena, x, nrhs, sb, SB, stream);
  return fwd ? launch_trsv<T, false, true, false>(fronts, flist, n_fronts, max_fs, arena, x, nrhs, sb, SB, stream)
             : launch_trsv<T, false, false, false>(fronts, flist, n_fronts, max_fs, arena, x, nrhs, sb, SB, stream);
}
}  // namespace

extern "C" {

int uhmk_fill_int(int* p, long long n, int value, hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, n, value);
  return ret(hipGetLastError());
}

int uhmk_lu_perm(const uhmk_front* fronts, const int* flist, int n_fronts, const int* piv, int* perm,
                 hipStream_t stream) {
  if (n_fronts <= 0) return 0;
  hipLaunchKernelGGL(k_lu_perm, dim3(n_fronts), dim3(64), 0, stream, fronts, flist, piv, perm);
  return ret(hipGetLastError());
}

int uhmk_solve_gather(const uhmk_front* fronts, const uhmk_item* items, int n_items, int lu, const int* src_beg,
                      const uhmk_source* srcs, const int* inv, const int* perm, double* x, const double* b, int nrhs,
                      hipStream_t stream) {
  if (n_items <= 0) return 0;
  if (lu)
    hipLaunchKernelGGL(k_sv_gather<true>, dim3(n_items), dim3(256), 0, stream, fronts, items, src_beg, srcs, inv,
                       perm, x, b, nrhs);
  else
    hipLaunchKernelGGL(k_sv_gather<false>, dim3(n_items), dim3(256), 0, stream, fronts, items, src_beg, srcs, inv,
                       perm, x, b, nrhs);
  return ret(hipGetLastError());
}


int uhmk_solve_trsv(const uhmk_front* fronts, const int* flist, int n_fronts, int max_fs, int chol, int fwd,
                    int sb, int sb_size, const void* arena, int prec, double* x, int nrhs, hipStream_t stream) {
  if (n_fronts <= 0 || max_fs <= 0) return 0;
  // chol = 2: Cholesky factors whose 64 x 64 diagonal blocks carry W = L^-1
  const int inv = chol == 2, ch = chol != 0;
  if (prec == 0)
    return trsv_any<double>(ch, fwd, inv, fronts, flist, n_fronts, max_fs, static_cast<const double*>(arena), x, nrhs,
                            sb, sb_size, stream);
  if (prec == 1)
    return trsv_any<float>(ch, fwd, inv, fronts, flist, n_fronts, max_fs, static_cast<const float*>(arena), x, nrhs,
                           sb, sb_size, stream);
  return -1006;
}

int uhmk_solve_gemv(const uhmk_front* fronts, const uhmk_item* items, int n_items, int cols, const void* arena,
                    int prec, double* x, double* dst, int nrhs, hipStream_t stream) {
  if (!dst) dst = x;
  if (n_items <= 0) return 0;
  if (prec != 0 && prec != 1) return -1006;
  if (!cols) {
    if (dst != x) return -1009;
    if (prec == 0)
      hipLaunchKernelGGL(k_sv_gemv_rows<double>, dim3(n_items), dim3(256), 0, stream, fronts, items,
                         static_cast<const double*>(arena), x, nrhs);
    else
      hipLaunchKernelGGL(k_sv_gemv_rows<float>, dim3(n_items), dim3(256), 0, stream, fronts, items,
                         static_cast<const float*>(arena), x, nrhs);
  } else {
    if (prec == 0)
      hipLaunchKernelGGL(k_sv_gemv_cols<double>, dim3(n_items), dim3(256), 0, stream, fronts, items,
                         static_cast<const double*>(arena), x, nrhs, dst);
    else
      hipLaunchKernelGGL(k_sv_gemv_cols<float>, dim3(n_items), dim3(256), 0, stream, fronts, items,
                         static_cast<const float*>(arena), x, nrhs, dst);
  }
  return ret(hipGetLastError());
}

int uhmk_solve_chol_step(const uhmk_front* fronts, const uhmk_item* items, int n_items, int fwd, int k0,
                         const void* arena, int prec, double* x, double* z, int nrhs, hipStream_t stream) {
  if (n_items <= 0) return 0;
  if (prec == 0) {
    if (fwd)
      hipLaunchKernelGGL(k_sv_chol_fwd_step<double>, dim3(n_items), dim3(256), 0, stream, fronts, items,
                         static_cast<const double*>(arena), x, z, nrhs, k0);
    else
      hipLaunchKernelGGL(k_sv_chol_bwd_step<double>, dim3(n_items), dim3(256), 0, stream, fronts, items,
                         static_cast<const double*>(arena), x, z, nrhs, k0);
  } else if (prec == 1) {
    if (fwd)
      hipLaunchKernelGGL(k_sv_chol_fwd_step<float>, dim3(n_items), dim3(256), 0, stream, fronts, items,
                         static_cast<const float*>(arena), x, z, nrhs, k0);
    else
      hipLaunchKernelGGL(k_sv_chol_bwd_step<float>, dim3(n_items), dim3(256), 0, stream, fronts, items,
                         static_cast<const float*>(arena), x, z, nrhs, k0);
  } else {
    return -1006;
  }
  return ret(hipGetLastError());
}

int uhmk_solve_chol_level(const uhmk_front* fronts, const uhmk_item* items, int n_items, int fwd,
                          const void* arena, int prec, double* x, double* z, int nrhs, int* err, hipStream_t stream) {
  if (n_items <= 0 || nrhs <= 0) return 0;
  if (nrhs > SOLVE_MAXR) return -1014;
  if (prec == 0) {
    if (fwd)
      hipLaunchKernelGGL(k_sv_chol_fwd_level<double>, dim3(n_items), dim3(256), 0, stream, fronts, items,
                         static_cast<const double*>(arena), x, z, nrhs, err);
    else
      hipLaunchKernelGGL(k_sv_chol_bwd_level<double>, dim3(n_items), dim3(256), 0, stream, fronts, items,
                         static_cast<const double*>(arena), x, z, nrhs, err);
  } else if (prec == 1) {
    if (fwd)
      hipLaunchKernelGGL(k_sv_chol_fwd_level<float>, dim3(n_items), dim3(256), 0, stream, fronts, items,
                         static_cast<const float*>(arena), x, z, nrhs, err);
    else
      hipLaunchKernelGGL(k_sv_chol_bwd_level<float>, dim3(n_items), dim3(256), 0, stream, fronts, items,
                         static_cast<const float*>(arena), x, z, nrhs, err);
  } else {
    return -1006;
  }
  return ret(hipGetLastError());
}

int uhmk_solve_mark(double* p, long long n, hipStream_t stream) {
  if (n <= 0) return 0;
  const long long nb = (n + 255) / 256;
  if (nb > 0x7fffffffLL) return -1011;
  hipLaunchKernelGGL(k_fill_mark, dim3((unsigned)nb), dim3(256), 0, stream, p, n);
  return ret(hipGetLastError());
}

int uhmk_solve_gather_cols(const uhmk_front* fronts, const uhmk_item* items, int n_items, int lu,
                           const int* src_beg, const uhmk_source* srcs, const int* inv, const int* perm, double* x,
                           const double* b, int nrhs, hipStream_t stream) {
  if (n_items <= 0 || nrhs <= 0) return 0;
  const long long ng = (nrhs + GCW - 1) / GCW, nb = ng * n_items;
  if (nb > 0x7fffffffLL) return -1011;
  if (lu)
    hipLaunchKernelGGL(k_sv_gather_cols<true>, dim3((unsigned)nb), dim3(64), 0, stream, fronts, items, src_beg, srcs,
                       inv, perm, x, b, nrhs, (int)ng);
  else
    hipLaunchKernelGGL(k_sv_gather_cols<false>, dim3((unsigned)nb), dim3(64), 0, stream, fronts, items, src_beg, srcs,
                       inv, perm, x, b, nrhs, (int)ng);
  return ret(hipGetLastError());
}

int uhmk_solve_branch_cols(const uhmk_front* fronts, const uhmk_item* items, int n_items, const int* emap, double* x,
                           int nrhs, hipStream_t stream) {
  if (n_items <= 0 || nrhs <= 0) return 0;
  const long long ng = (nrhs + GCW - 1) / GCW, nb = ng * n_items;
  if (nb > 0x7fffffffLL) return -1011;
  hipLaunchKernelGGL(k_sv_branch_cols, dim3((unsigned)nb), dim3(64), 0, stream, fronts, items, emap, x, nrhs,
                     (int)ng);
  return ret(hipGetLastError());
}

int uhmk_solve_branch(const uhmk_front* fronts, const uhmk_item* items, int n_items, const int* emap, double* x,
                      int nrhs, hipStream_t stream) {
  if (n_items <= 0) return 0;
  hipLaunchKernelGGL(k_sv_branch, dim3(n_items), dim3(256), 0, stream, fronts, items, emap, x, nrhs);
  return ret(hipGetLastError());
}

int uhmk_generate_elements(const uhmk_front* fronts, const uhmk_item* items, int n_items, int max_m,
                           const int* lmap, unsigned long long seed, double alpha, double beta,
                           int scale_kind, double p0, double p1, double* stage, double* b, int nrhs,
                           int with_lhs, int with_rhs, hipStream_t stream) {
  if (n_items <= 0) return 0;
  const int ldm = max_m | 1;
  const size_t sh = with_lhs ? (size_t)ldm * max_m * sizeof(double) : 0;
  if (sh > 160 * 1024) return -1003;
  hipLaunchKernelGGL(k_gen_elements, dim3(n_items), dim3(256), sh, stream, fronts, items, lmap, seed, alpha, beta,
                     scale_kind, p0, p1, stage, b, nrhs, with_lhs, with_rhs);
  return ret(hipGetLastError());
}

int uhmk_csr_values(long long nnz, const long long* cptr, const long long* cidx, const double* stage, double* val,
                    hipStream_t stream) {
  if (nnz <= 0) return 0;
  hipLaunchKernelGGL(k_csr_values, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, stream, nnz, cptr, cidx, stage,
                     val);
  return ret(hipGetLastError());
}

int uhmk_gather_rhs(long long nrow, const long long* bptr, const long long* bidx, const int* bn, const double* b,
                    double* bg, int col, hipStream_t stream) {
  if (nrow <= 0) return 0;
  hipLaunchKernelGGL(k_gather_rhs, dim3((unsigned)((nrow + 255) / 256)), dim3(256), 0, stream, nrow, bptr, bidx, bn,
                     b, bg, col);
  return ret(hipGetLastError());
}

}  // extern "C"
