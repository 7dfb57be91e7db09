// Device helpers shared by the dense front kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

namespace uhmd {

// p[idx] when ok, else zero, without a branch: the address falls back to
// p[0] (which must be a valid address) and the value is selected after the
// load, so that a group of such loads issues back to back (a conditional
// load followed by a conversion would otherwise wait inside its branch).
template <typename T>
__device__ __forceinline__ double ldz(const T* p, long long idx, bool ok) {
  const T v = p[ok ? idx : 0];
  return ok ? (double)v : 0.0;
}
// The same as a value of type TM, converted after the load.  FP64 to FP64
// keeps the predicated form (`ok ? p[idx] : 0`, one exec-mask branch per
// load), which measured 4-5 % faster in the FP64 MFMA update; every other
// pair loads from the clamped address, which measured 4 % faster in the FP32
// update (the loads issue back to back).
template <typename TM, typename T>
__device__ __forceinline__ TM ldz_as(const T* p, long long idx, bool ok) {
  if constexpr (sizeof(TM) == 8 && sizeof(T) == 8) {
    return ok ? (TM)p[idx] : (TM)0;
  } else {
    const T v = p[ok ? idx : 0];
    return ok ? (TM)v : (TM)0;
  }
}

// Item of workgroup b of a launch over n items, the items dealt out to the
// eight XCDs in contiguous runs (workgroup b runs on XCD b % 8, and XCD x
// takes items [x n / 8, (x + 1) n / 8) in order): tiles that share panels
// sit next to each other in the item list, so each XCD's L2 keeps its own
// panels instead of all eight fetching every panel.  A bijection on [0, n).
__device__ __forceinline__ int xcd_item(int b, int n) {
  const int x = b & 7, l = b >> 3, q = n >> 3, r = n & 7;
  return x < r ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
}

// Broadcast of a double from lane (l & ~3) + KQ to the four lanes of its
// quad (DPP quad_perm, no LDS round trip).
template <int KQ>
__device__ __forceinline__ double quad_bcast(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), KQ * 85, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), KQ * 85, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Value of v in lane L (uniform result, v_readlane).
template <int L>
__device__ __forceinline__ double readlane_d(double v) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), L),
                          __builtin_amdgcn_readlane(__double2loint(v), L));
}

// 1 / sqrt(d) from v_rsq_f64 refined by two Newton steps (full FP64
// accuracy for positive normal d; NaN / Inf otherwise).
__device__ __forceinline__ double rsqrt_d(double d) {
  double r = __builtin_amdgcn_rsq(d);
  const double h = 0.5 * d;
  r = r * fma(-h * r, r, 1.5);
  r = r * fma(-h * r, r, 1.5);
  return r;
}

// Triangular solve on 64-long vectors, 16 vectors per wavefront, quad
// layout: lane l = 4 v + q holds entries [16 q, 16 q + 16) of vector v in
// x[0..16).  For k = 0 .. kb-1 in order:
//     x_k <- x_k * rdiag[k]                 (skipped when UNIT)
//     x_c <- x_c - x_k * T(c, k)   for k < c < kb
// with T(c, k) = s[c * ld + k] (TRANS = false) or s[k * ld + c] (TRANS = true)
// read from LDS.  This is the right-looking (column) form of
//   * X <- X L^-T          rows of X, T = L    (Cholesky panel rows)
//   * X <- X U^-1          rows of X, T = U^T  (LU rows of ABL, TRANS)
//   * x <- L^-1 x, unit    columns,   T = L    (LU U strip)
// Each vector is independent; the owner of x_k broadcasts it inside its
// quad with DPP.  The 64 steps are expanded at compile time so every
// register index is static.  The caller stores T with zeros outside the
// strict triangle that is used (and outside kb), so every update is an
// unconditional FMA.
// ST: element type of s (FP32 triangles of FP32 fronts take half the LDS
// reads; each entry is converted to FP64, exactly, before its FMA).
template <int K, bool UNIT, bool TRANS, bool FULL, typename ST = double>
struct TriStep {
  __device__ __forceinline__ static void run(double (&x)[16], const ST* __restrict__ s, int ld,
                                             const double* __restrict__ rdiag, int kb, int q) {
    if (FULL || K < kb) {
      constexpr int KQ = K >> 4, KR = K & 15;
      if (!UNIT) {
        const double xo = x[KR];
        x[KR] = (q == KQ) ? xo * rdiag[K] : xo;
      }
      const double xk = quad_bcast<KQ>(x[KR]);
#pragma unroll
      for (int cr = 0; cr < 16; ++cr) {
        const int c = 16 * q + cr;
        x[cr] = fma(-xk, (double)(TRANS ? s[K * ld + c] : s[c * ld + K]), x[cr]);
      }
      TriStep<K + 1, UNIT, TRANS, FULL, ST>::run(x, s, ld, rdiag, kb, q);
    }
  }
};
template <bool UNIT, bool TRANS, bool FULL, typename ST>
struct TriStep<64, UNIT, TRANS, FULL, ST> {
  __device__ __forceinline__ static void run(double (&)[16], const ST*, int, const double*, int, int) {}
};

// s must hold T(c, k) for k < c < kb and zero everywhere else in the
// 64 x 64 block (the diagonal lives in rdiag).
template <bool UNIT, bool TRANS, typename ST>
__device__ __forceinline__ void tri_solve64(double (&x)[16], const ST* __restrict__ s, int ld,
                                            const double* __restrict__ rdiag, int kb, int lane) {
  if (kb == 64) TriStep<0, UNIT, TRANS, true, ST>::run(x, s, ld, rdiag, kb, lane & 3);
  else TriStep<0, UNIT, TRANS, false, ST>::run(x, s, ld, rdiag, kb, lane & 3);
}

// Fill a 64 x 64 LDS block (row stride ld) with the strict lower (LOWER)
// or strict upper triangle of the kb x kb block at A (leading dimension n),
// zeros elsewhere, and rdiag[k] = 1 / A[k][k] (k < kb).  NT threads; every
// thread's loads are issued before its LDS stores so they overlap.  ST, the
// LDS element type, may be FP32 for an FP32 front (its entries are exact).
template <bool LOWER, int NT, typename T, typename ST>
__device__ __forceinline__ void load_tri64(ST* __restrict__ s, int ld, double* __restrict__ rdiag,
                                           const T* __restrict__ A, long long n, int kb, int tid) {
  constexpr int PER = 64 * 64 / NT;
  double v[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int idx = tid + NT * u, r = idx & 63, c = idx >> 6;
    v[u] = ldz(A, r + c * n, r < kb && c < kb && (LOWER ? r >= c : r <= c));
  }
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int idx = tid + NT * u, r = idx & 63, c = idx >> 6;
    s[r * ld + c] = (r == c) ? (ST)0 : (ST)v[u];
    if (r == c && r < kb) rdiag[r] = 1.0 / v[u];
  }
}

// Right-looking Cholesky of a 64 x 64 block by four wavefronts: lane i
// holds row i, wavefront w holds columns j = w + 4 u (u = 0..15) in a[u].
// Step K: the owner wavefront of column K takes the pivot from lane K with
// v_readlane, scales the column (1 / sqrt from v_rsq_f64 and two Newton
// steps) and publishes it in LDS (two alternating buffers, so one barrier
// per step suffices); every wavefront then updates its columns right of K.
// Entries above the diagonal are zero on return.  first <- K + 1 of the
// first non-positive pivot (owner wavefronts only; combine across waves).
template <int K>
struct CholColStep {
  __device__ __forceinline__ static void run(double (&a)[16], int i, int w, int kb, int& first,
                                             double (*sc)[64]) {
    if (K < kb) {
      constexpr int U = K >> 2, WK = K & 3;
      if (w == WK) {
        const double d = readlane_d<K>(a[U]);
        if (!(d > 0.0) && first == 0) first = K + 1;
        const double r = rsqrt_d(d);
        a[U] = (i == K) ? d * r : (i > K ? a[U] * r : 0.0);
        sc[K & 1][i] = a[U];
      }
      __syncthreads();
      const double lik = sc[K & 1][i];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int j = w + 4 * u;
        if (j > K) a[u] = fma(-lik, sc[K & 1][j], a[u]);
      }
      CholColStep<K + 1>::run(a, i, w, kb, first, sc);
    }
  }
};
template <>
struct CholColStep<64> {
  __device__ __forceinline__ static void run(double (&)[16], int, int, int, int&, double (*)[64]) {}
};

}  // namespace uhmd
