// Front factorization kernels for gfx950 (FP64 path).
//
// One launch covers every front of a tree level.  Work items (uhmk_item)
// are built once per sparsity pattern on the host, so the whole level
// schedule is a fixed sequence of launches (and can be captured in a graph).
//
// Reference operations restated (semantics, not code):
//   extend-add   Helper_::_merge_A + Matrix_FLA_::merge
//                UHM/include/uhm/matrix/uhm/helper.hxx:191-222,
//                UHM/src/matrix/uhm/fla/matrix.cxx:284-303
//   LU front     lu_piv_hier, UHM/src/matrix/uhm/fla/lu_piv/decompose.cxx:124-153
//                linal::dense::lu_piv, LINAL/src/dense/lu_piv.cxx:47-111
//                apply_pivots, LINAL/src/dense/pivot.cxx:46-91
//   Chol front   chol_hier, UHM/src/matrix/uhm/fla/chol/decompose.cxx:85-98
//                linal::dense::chol, LINAL/src/dense/chol.cxx:46-92
//   trsm / gemm / syrk  LINAL/src/dense/{trsm,gemm,syrk}.cxx
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "assemble_device.hpp"
#include "dense_device.hpp"
#include "uhmk.h"

namespace {

constexpr int TILE = 64;

// ------------------------------------------------------------------------
// Extend-add in gather form.  Every parent entry is produced exactly once:
//   F[i,j] = ((0 + stage[i,j]) + C1[inv1[i],inv1[j]]) + C2[...] ...
// summing the sources in the reference's merge order (own copy-in first,
// then children in insertion order) into a zero start, which is the exact
// operation sequence of op_merge on a freshly zeroed front
// (UHM/src/operation/element.cxx:135-185).  Item: f, a = row tile start,
// b = column tile start.  The sum is formed in FP64 whatever the storage
// types: T of the assembled front, FP64 or FP32 per child (uhmd::asm_tile).
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_assemble(const uhmk_front* __restrict__ F,
                                                  const uhmk_item* __restrict__ items,
                                                  const int* __restrict__ src_beg,
                                                  const uhmk_source* __restrict__ srcs,
                                                  const int* __restrict__ inv,
                                                  double* __restrict__ arena64,
                                                  float* __restrict__ arena32,
                                                  const double* __restrict__ stage) {
  const uhmk_item it = items[blockIdx.x];
  __shared__ uhmd::AsmSources S;
  uhmd::asm_sources(S, F, it.f, src_beg, srcs);
  __syncthreads();
  uhmd::asm_tile<T>(S, F[it.f].n, F[it.f].off, it.a, it.b, inv, arena64, arena32, stage);
}

// Copy-in blocks of one (front, side) item added in order into the front's
// stage (side 0) or right-hand side block (side 1), zeroed first when the
// item says so.
__global__ __launch_bounds__(256) void k_copy_in(const uhmk_front* __restrict__ F, const uhmk_item* __restrict__ items,
                                                 const uhmk_cblock* __restrict__ blocks,
                                                 const double* __restrict__ vals, const int* __restrict__ maps,
                                                 double* __restrict__ stage, double* __restrict__ b, int nrhs) {
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, side = it.c & 1;
  if (side == 0 && fr.soff < 0) return;
  double* __restrict__ T = side ? b + fr.xoff : stage + fr.soff;
  const long long len = (long long)n * (side ? nrhs : n);
  if (it.c & 2) {
    for (long long i = threadIdx.x; i < len; i += 256) T[i] = 0.0;
    __syncthreads();
  }
  for (int q = 0; q < it.b; ++q) {
    const uhmk_cblock c = blocks[it.a + q];
    const int* __restrict__ mp = maps + c.moff;
    const double* __restrict__ K = vals + c.koff;
    const long long cnt = (long long)c.m * c.ncol;
    for (long long idx = threadIdx.x; idx < cnt; idx += 256) {
      const int i = (int)(idx % c.m), j = (int)(idx / c.m);
      const long long col = side ? j : mp[j];
      T[mp[i] + col * n] += K[idx];
    }
    __syncthreads();
  }
}

// dst[i] = (TD)src[i], i < count (a front's entries into another arena)
template <typename TS, typename TD>
__global__ __launch_bounds__(256) void k_copy_cast(const TS* __restrict__ src, TD* __restrict__ dst, long long count) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < count) dst[i] = (TD)src[i];
}

// ------------------------------------------------------------------------
// Partial LU panel on rows [c0, fs) x columns [c0, c0+kb) of each front,
// pivots searched in rows [c0, fs) only (the reference restricts pivoting to
// ATL).  One workgroup (512 threads) per front; the panel is processed in
// inner blocks of IB columns held in registers, R rows per thread.  Pivot
// choice is the first row of maximal |a| (libflame / LAPACK i?amax), found
// with a wavefront shuffle reduction and a cross-wave pass in LDS.
// Swaps are applied to the whole outer panel here; the row kernel applies
// them to every other column.
// ------------------------------------------------------------------------
constexpr int PIB = 8;

// Multiplier a / pv of a row below the pivot pv, formed as LAPACK dgetf2
// forms it: a times the reciprocal rinv = 1 / pv when |pv| is at least the
// smallest normal number, by division when pv is subnormal, and a itself
// when pv is zero (the zero pivot is reported instead).
__device__ __forceinline__ double lu_mult(double a, double pv, double rinv) {
  if (fabs(pv) >= 2.2250738585072014e-308) return a * rinv;
  return pv != 0.0 ? a / pv : a;
}

constexpr int PP_STRIDE = 1 + 4 * TILE;  // ints per front in the panel-permutation buffer
constexpr int PP_MAXROWS = 6144;

// Net row permutation of a finished panel for the row kernel: rows whose
// content changed (the kb strip rows in order, then the distinct swap
// targets below the strip in order of first use) and, per listed row, the
// list index of the row whose old value lands there.  s_piv: the panel's
// relative pivots; m = panel rows (fs - c0).
template <int PNT>
__device__ __forceinline__ void panel_permutation(const int* s_piv, int kb, int m, int c0, int* __restrict__ pp,
                                                  int tid) {
  __shared__ int s_cont[PP_MAXROWS];
  __shared__ int s_tgt[TILE];
  __shared__ int s_ncnt;
  for (int i = tid; i < m; i += PNT) s_cont[i] = i;
  __syncthreads();
  if (tid == 0)
    for (int k = 0; k < kb; ++k) {
      const int d = s_piv[k];
      if (d) {
        const int t = s_cont[k];
        s_cont[k] = s_cont[k + d];
        s_cont[k + d] = t;
      }
    }
  if (tid < 64) {
    const int k = tid;
    int t = -1;
    bool first = false;
    if (k < kb) {
      t = k + s_piv[k];
      first = t >= kb;
      for (int k2 = 0; k2 < k && first; ++k2)
        if (k2 + s_piv[k2] == t) first = false;
    }
    const unsigned long long mask = __ballot(first);
    if (first) s_tgt[__popcll(mask & ((1ull << k) - 1ull))] = t;
    if (k == 0) s_ncnt = kb + __popcll(mask);
  }
  __syncthreads();
  const int cnt = s_ncnt;
  for (int k = tid; k < cnt; k += PNT) {
    const int lr = (k < kb) ? k : s_tgt[k - kb];
    const int src = s_cont[lr];
    int si = (src < kb) ? src : -1;
    for (int u = 0; u < cnt - kb && si < 0; ++u)
      if (s_tgt[u] == src) si = kb + u;
    pp[1 + k] = c0 + lr;
    pp[1 + 2 * TILE + k] = si;
  }
  if (tid == 0) pp[0] = cnt;
}

template <int R, typename T, int PNT>
__global__ __launch_bounds__(PNT) void k_lu_panel(const uhmk_front* __restrict__ F,
                                                  const int* __restrict__ flist, int c0, int nb,
                                                  T* __restrict__ arena, int* __restrict__ piv,
                                                  int* __restrict__ info, int* __restrict__ pperm) {
  const int f = flist[blockIdx.x];
  const uhmk_front fr = F[f];
  const int n = fr.n, fs = fr.fs;
  const int kb = min(nb, fs - c0);
  T* __restrict__ A = arena + fr.off;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  constexpr int PNW = PNT / 64;

  // per-column candidate slots, double-buffered by column parity: one
  // barrier per column publishes every wavefront's best row (value, index
  // and its inner-block values) and the current row's values
  __shared__ double s_cv[2][PNW];
  __shared__ int s_ci[2][PNW];
  __shared__ double s_crow[2][PNW][PIB];
  __shared__ double s_cur[2][PIB];
  __shared__ int s_piv[TILE];
  __shared__ double s_L[PIB][PIB];
  __shared__ double s_U[PIB][TILE];
  __shared__ int s_prow[2 * PIB], s_psrc[2 * PIB];
  __shared__ int s_pcnt;
  __shared__ double s_tmp[2 * PIB][TILE];

  int first_zero = 0;  // thread 0: first zero-pivot row + 1 of this panel
  for (int ib = 0; ib < kb; ib += PIB) {
    const int ibw = min(PIB, kb - ib);
    const int rb = c0 + ib;
    double v[R][PIB];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = rb + tid + PNT * q;
#pragma unroll
      for (int j = 0; j < PIB; ++j) v[q][j] = uhmd::ldz(A, r + (long long)(rb + j) * n, r < fs && j < ibw);
    }
#pragma unroll
    for (int j = 0; j < PIB; ++j) {
      if (j < ibw) {
        const int cr = rb + j, buf = j & 1;
        // best row of this thread, then of the wavefront (first maximal |a|)
        double bv = -1.0;
        int bi = 0x7fffffff;
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int r = rb + tid + PNT * q;
          if (r >= cr && r < fs) {
            const double a = fabs(v[q][j]);
            if (a > bv) {
              bv = a;
              bi = r;
            }
          }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          const double ov = __shfl_xor(bv, o);
          const int oi = __shfl_xor(bi, o);
          if (ov > bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
          }
        }
        if (lane == 0) {
          s_cv[buf][wid] = bv;
          s_ci[buf][wid] = bi;
        }
        // the owner of the wavefront's best row publishes its values, the
        // owner of the current row (thread j, first row) publishes its own
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int r = rb + tid + PNT * q;
          if (r == bi)
#pragma unroll
            for (int jj = 0; jj < PIB; ++jj) s_crow[buf][wid][jj] = v[q][jj];
        }
        if (tid == j)
#pragma unroll
          for (int jj = 0; jj < PIB; ++jj) s_cur[buf][jj] = v[0][jj];
        __syncthreads();
        // every thread reduces the wavefront candidates itself
        double b2 = s_cv[buf][0];
        int i2 = s_ci[buf][0], wb = 0;
#pragma unroll
        for (int w = 1; w < PNW; ++w) {
          const double cv = s_cv[buf][w];
          const int ci = s_ci[buf][w];
          if (cv > b2 || (cv == b2 && ci < i2)) {
            b2 = cv;
            i2 = ci;
            wb = w;
          }
        }
        // a column of NaNs has no maximum: keep the current row
        const bool keep = (i2 < cr || i2 >= fs);
        const int pi = keep ? cr : i2;
        const double* __restrict__ prow = keep ? s_cur[buf] : s_crow[buf][wb];
        const double* __restrict__ crow = s_cur[buf];
        if (pi != cr) {
#pragma unroll
          for (int q = 0; q < R; ++q) {
            const int r = rb + tid + PNT * q;
            if (r == pi)
#pragma unroll
              for (int jj = 0; jj < PIB; ++jj) v[q][jj] = crow[jj];
            if (r == cr)
#pragma unroll
              for (int jj = 0; jj < PIB; ++jj) v[q][jj] = prow[jj];
          }
        }
        const double pv = prow[j], rinv = 1.0 / pv;
        if (tid == 0) {
          s_piv[ib + j] = pi - cr;
          if (!(pv != 0.0) && first_zero == 0) first_zero = cr + 1;
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int r = rb + tid + PNT * q;
          if (r > cr && r < fs) {
            const double l = lu_mult(v[q][j], pv, rinv);
            v[q][j] = l;
#pragma unroll
            for (int kk = j + 1; kk < PIB; ++kk) v[q][kk] = fma(-l, prow[kk], v[q][kk]);
          }
        }
      }
    }
    __syncthreads();
    // store the factored inner block
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = rb + tid + PNT * q;
      if (r < fs)
#pragma unroll
        for (int jj = 0; jj < PIB; ++jj)
          if (jj < ibw) A[r + (long long)(rb + jj) * n] = (T)v[q][jj];
    }
    __syncthreads();
    // Net row permutation of this inner block's swaps (rows listed once;
    // s_psrc[k] = list index of the row whose old value lands on s_prow[k]).
    if (tid == 0) {
      int cnt = 0;
      int content[2 * PIB];
      for (int s = 0; s < ibw; ++s) {
        const int ra = rb + s, rb2 = rb + s + s_piv[ib + s];
        int ia = -1, ib2 = -1;
        for (int k = 0; k < cnt; ++k) {
          if (s_prow[k] == ra) ia = k;
          if (s_prow[k] == rb2) ib2 = k;
        }
        if (ia < 0) {
          s_prow[cnt] = ra;
          content[cnt] = ra;
          ia = cnt++;
        }
        if (ib2 < 0) {
          s_prow[cnt] = rb2;
          content[cnt] = rb2;
          ib2 = cnt++;
        }
        const int t = content[ia];
        content[ia] = content[ib2];
        content[ib2] = t;
      }
      for (int k = 0; k < cnt; ++k) {
        int src = k;
        for (int k2 = 0; k2 < cnt; ++k2)
          if (s_prow[k2] == content[k]) src = k2;
        s_psrc[k] = src;
      }
      s_pcnt = cnt;
    }
    __syncthreads();
    const int pc = s_pcnt;
    const int nother = kb - ibw;
    for (int idx = tid; idx < nother * pc; idx += PNT) {
      const int t = idx / pc, k = idx - t * pc;
      const int col = (t < ib) ? (c0 + t) : (c0 + t + ibw);
      s_tmp[k][t] = (double)A[s_prow[k] + (long long)col * n];
    }
    __syncthreads();
    for (int idx = tid; idx < nother * pc; idx += PNT) {
      const int t = idx / pc, k = idx - t * pc;
      const int col = (t < ib) ? (c0 + t) : (c0 + t + ibw);
      A[s_prow[k] + (long long)col * n] = (T)s_tmp[s_psrc[k]][t];
    }
    const int nr = kb - ib - ibw;
    if (tid < PIB) {
#pragma unroll
      for (int jj = 0; jj < PIB; ++jj) s_L[tid][jj] = (tid < ibw) ? v[0][jj] : 0.0;
    }
    __syncthreads();
    if (nr > 0) {
      // U strip of the columns right of this inner block (inside the panel)
      for (int t = tid; t < nr; t += PNT) {
        const int col = rb + ibw + t;
        double u[PIB];
#pragma unroll
        for (int jj = 0; jj < PIB; ++jj) u[jj] = uhmd::ldz(A, rb + jj + (long long)col * n, jj < ibw);
#pragma unroll
        for (int jj = 1; jj < PIB; ++jj)
#pragma unroll
          for (int kk = 0; kk < jj; ++kk) u[jj] -= s_L[jj][kk] * u[kk];
#pragma unroll
        for (int jj = 0; jj < PIB; ++jj) {
          if (jj < ibw) A[rb + jj + (long long)col * n] = (T)u[jj];
          s_U[jj][t] = (jj < ibw) ? u[jj] : 0.0;
        }
      }
      __syncthreads();
      // rank-ibw update of the rows below for those columns: per step every
      // thread loads CW columns of all its rows before computing, so the
      // loads of a step are in flight together
      constexpr int CW = (R <= 2) ? 8 : ((R <= 4) ? 4 : 2);
#pragma unroll 1
      for (int t0 = 0; t0 < nr; t0 += CW) {
        double a[R][CW];
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int r = rb + tid + PNT * q;
          const bool rv = r >= rb + ibw && r < fs;
#pragma unroll
          for (int u = 0; u < CW; ++u)
            a[q][u] = uhmd::ldz(A, r + (long long)(rb + ibw + t0 + u) * n, rv && t0 + u < nr);
        }
#pragma unroll
        for (int kk = 0; kk < PIB; ++kk) {
          double uu[CW];
#pragma unroll
          for (int u = 0; u < CW; ++u) uu[u] = s_U[kk][(t0 + u) & (TILE - 1)];
#pragma unroll
          for (int q = 0; q < R; ++q)
#pragma unroll
            for (int u = 0; u < CW; ++u) a[q][u] = fma(-v[q][kk], uu[u], a[q][u]);
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int r = rb + tid + PNT * q;
          if (r >= rb + ibw && r < fs)
#pragma unroll
            for (int u = 0; u < CW; ++u)
              if (t0 + u < nr) A[r + (long long)(rb + ibw + t0 + u) * n] = (T)a[q][u];
        }
      }
      __syncthreads();
    }
  }
  for (int k = tid; k < kb; k += PNT) piv[fr.poff + c0 + k] = s_piv[k];
  if (tid == 0 && first_zero && info[f] == 0) info[f] = first_zero;
  panel_permutation<PNT>(s_piv, kb, fs - c0, c0, pperm + (long long)f * PP_STRIDE, tid);
}

// ------------------------------------------------------------------------
// After a panel: row swaps for every column outside the panel, the U strip
// U12 = L11^-1 * A12 for columns right of the panel (unit lower TRSM), and
// the ABL rows of the panel L21 = A21 * U11^-1 (rows [fs, n), right upper
// TRSM).  Item: f, a = kind (0 left columns, 1 ABL rows, 2 right columns),
// b = first column / row of the 64-wide chunk.
// ------------------------------------------------------------------------
constexpr int LDP = TILE + 1;

// ------------------------------------------------------------------------
// After a panel: row swaps for every column outside the panel
// (apply_pivots), the U strip U12 = L11^-1 A12 for the columns right of the
// panel, and the ABL rows of the panel L21 = A21 U11^-1 (rows [fs, n)).
// Item: f, a = kind (0 left columns, 1 ABL rows, 2 right columns),
// b = first column / row of its chunk of NT / 4 (16, 32 or 64).  The
// triangular solves use 4 lanes per vector (tri_solve64), NT / 4 vectors
// per workgroup; narrow chunks spread one panel's row work over more CUs.
// ------------------------------------------------------------------------

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_lu_rows(const uhmk_front* __restrict__ F,
                                                 const uhmk_item* __restrict__ items, int c0, int nb,
                                                 T* __restrict__ arena, const int* __restrict__ pperm) {
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs;
  const int kb = min(nb, fs - c0);
  T* __restrict__ A = arena + fr.off;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, q = lane & 3;
  constexpr int CW = NT / 4;  // columns (rows for ABL items) per item: four lanes each in the solves
  __shared__ double s_X[2 * TILE][CW + 1];  // rows touched by the swaps (row stride CW + 1: the
                                            // gather's consecutive rows on different banks)
  // triangular panel block, in the front's storage type (FP32 fronts: half
  // the LDS reads of the triangular solves, and a third workgroup per CU)
  typedef typename std::conditional<sizeof(T) == 4, float, double>::type ST;
  __shared__ ST s_T[TILE * LDP];
  __shared__ double s_rd[TILE];
  __shared__ int s_prow[2 * TILE], s_psrc[2 * TILE];
  __shared__ int s_cnt;

  if (it.a == 1) {
    uhmd::load_tri64<false, NT>(s_T, LDP, s_rd, A + c0 + (long long)c0 * n, n, kb, tid);
    __syncthreads();
    const int row = it.b + 16 * w + (lane >> 2);
    const bool live = row < n;
    double x[16];
#pragma unroll
    for (int cr = 0; cr < 16; ++cr) {
      const int c = 16 * q + cr;
      x[cr] = uhmd::ldz(A, row + (long long)(c0 + c) * n, live && c < kb);
    }
    uhmd::tri_solve64<false, true>(x, s_T, LDP, s_rd, kb, lane);
    if (live)
#pragma unroll
      for (int cr = 0; cr < 16; ++cr) {
        const int c = 16 * q + cr;
        if (c < kb) A[row + (long long)(c0 + c) * n] = (T)x[cr];
      }
    return;
  }

  const int j0 = it.b;
  const int width = min(CW, (it.a == 0 ? c0 : n) - j0);
  {
    const int* __restrict__ pp = pperm + (long long)it.f * PP_STRIDE;
    if (tid == 0) s_cnt = pp[0];
    if (tid < 2 * TILE) {
      s_prow[tid] = pp[1 + tid];
      s_psrc[tid] = pp[1 + 2 * TILE + tid];
    }
  }
  __syncthreads();
  const int pc = s_cnt;
  for (int idx = tid; idx < width * pc; idx += NT) {
    const int k = idx % pc, t = idx / pc;
    s_X[k][t] = (double)A[s_prow[k] + (long long)(j0 + t) * n];
  }
  __syncthreads();
  if (it.a == 0) {
    for (int idx = tid; idx < width * pc; idx += NT) {
      const int k = idx % pc, t = idx / pc;
      A[s_prow[k] + (long long)(j0 + t) * n] = (T)s_X[s_psrc[k]][t];
    }
    return;
  }
  // right columns: rows moved below the strip go back now; the strip is
  // solved in registers, one column per 4 lanes.
  for (int idx = tid; idx < width * (pc - kb); idx += NT) {
    const int k = kb + idx % (pc - kb), t = idx / (pc - kb);
    A[s_prow[k] + (long long)(j0 + t) * n] = (T)s_X[s_psrc[k]][t];
  }
  uhmd::load_tri64<true, NT>(s_T, LDP, s_rd, A + c0 + (long long)c0 * n, n, kb, tid);
  __syncthreads();
  const int t = 16 * w + (lane >> 2);
  double x[16];
#pragma unroll
  for (int cr = 0; cr < 16; ++cr) {
    const int k = 16 * q + cr;
    x[cr] = (t < width && k < kb) ? s_X[s_psrc[k]][t] : 0.0;
  }
  uhmd::tri_solve64<true, false>(x, s_T, LDP, s_rd, kb, lane);
  if (t < width)
#pragma unroll
    for (int cr = 0; cr < 16; ++cr) {
      const int k = 16 * q + cr;
      if (k < kb) A[(c0 + k) + (long long)(j0 + t) * n] = (T)x[cr];
    }
}

inline int ret(hipError_t e) { return e == hipSuccess ? 0 : -(int)e; }

// ------------------------------------------------------------------------
// The same panel with the whole panel in registers, for panels of at most
// PNT <= 512 pivot-candidate rows: thread t holds panel row c0 + t (all kb
// columns), so no panel data travels through memory between the column
// steps.  Rows are not moved: each thread tracks its row's current position
// and the rows are stored at their final positions at the end.  Per column
// two barriers: one publishes the wavefronts' pivot candidates, one the
// pivot row (written to LDS by its owner, double-buffered by column
// parity).  The columns run in groups of 8 (group index at run time,
// column inside the group and every register index at compile time).
// Pivot rule and arithmetic are those of k_lu_panel (first maximal |a| in
// current order, multipliers as lu_mult forms them, rank-1 FMA updates in
// pivot order), so both kernels give the same bits.
// ------------------------------------------------------------------------
template <typename T, int PNT>
__global__ __launch_bounds__(PNT) void k_lu_panel_reg(const uhmk_front* __restrict__ F,
                                                      const int* __restrict__ flist, int c0, int nb,
                                                      T* __restrict__ arena, int* __restrict__ piv,
                                                      int* __restrict__ info, int* __restrict__ pperm) {
  constexpr int NW = PNT / 64;
  __shared__ double s_cv[2][NW];
  __shared__ int s_ci[2][NW];
  __shared__ double s_row[2][TILE];
  __shared__ int s_piv[TILE];
  const int f = flist[blockIdx.x];
  const uhmk_front fr = F[f];
  const int n = fr.n, fs = fr.fs;
  const int kb = min(nb, fs - c0), m = fs - c0;
  T* __restrict__ A = arena + fr.off + c0 + (long long)c0 * n;  // panel origin (c0, c0)
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const bool live = tid < m;
  double a[TILE];
#pragma unroll
  for (int j = 0; j < TILE; ++j) a[j] = uhmd::ldz(A, tid + (long long)j * n, live && j < kb);
  int pos = tid, zero = 0;
  for (int g = 0; 8 * g < kb; ++g) {
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      const int K = 8 * g + kk;
      if (K >= kb) break;
      const int buf = kk & 1;
      double ak = 0.0;
#pragma unroll
      for (int h = 0; h < 8; ++h)
        if (h == g) ak = a[8 * h + kk];
      // first maximal |a(., K)| over the rows at positions [K, m)
      double cv = -1.0;
      int ck = 0x7fffffff;
      if (live && pos >= K) {
        const double v = fabs(ak);
        if (v > cv) {
          cv = v;
          ck = pos;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(cv, o);
        const int ok = __shfl_xor(ck, o);
        if (ov > cv || (ov == cv && ok < ck)) {
          cv = ov;
          ck = ok;
        }
      }
      if (lane == 0) {
        s_cv[buf][w] = cv;
        s_ci[buf][w] = ck;
      }
      __syncthreads();
      double bv = s_cv[buf][0];
      int bk = s_ci[buf][0];
#pragma unroll
      for (int q = 1; q < NW; ++q) {
        const double ov = s_cv[buf][q];
        const int ok = s_ci[buf][q];
        if (ov > bv || (ov == bv && ok < bk)) {
          bv = ov;
          bk = ok;
        }
      }
      // a column of NaNs has no maximum: keep the current row
      const int pp = (bk < K || bk >= m) ? K : bk;
      if (live && pos == pp) {
#pragma unroll
        for (int h = 0; h < 8; ++h) {
          if (h < g) continue;
#pragma unroll
          for (int jj = 0; jj < 8; ++jj)
            if (8 * h + jj < kb) s_row[buf][8 * h + jj] = a[8 * h + jj];
        }
      }
      if (live) {
        if (pos == pp) pos = K;
        else if (pos == K) pos = pp;
      }
      if (tid == 0) s_piv[K] = pp - K;
      __syncthreads();
      const double pv = s_row[buf][K], rinv = 1.0 / pv;
      if (tid == 0 && !(pv != 0.0) && zero == 0) zero = K + 1;
      if (live && pos > K) {
        const double l = lu_mult(ak, pv, rinv);
#pragma unroll
        for (int h = 0; h < 8; ++h) {
          if (h < g) continue;
          if (h == g) a[8 * h + kk] = l;
#pragma unroll
          for (int jj = 0; jj < 8; ++jj) {
            const int j = 8 * h + jj;
            if ((h > g || jj > kk) && j < kb) a[j] = fma(-l, s_row[buf][j], a[j]);
          }
        }
      }
    }
  }
  if (live)
#pragma unroll
    for (int j = 0; j < TILE; ++j)
      if (j < kb) A[pos + (long long)j * n] = (T)a[j];
  __syncthreads();
  for (int k = tid; k < kb; k += PNT) piv[fr.poff + c0 + k] = s_piv[k];
  if (tid == 0 && zero && info[f] == 0) info[f] = c0 + zero;
  panel_permutation<PNT>(s_piv, kb, m, c0, pperm + (long long)f * PP_STRIDE, tid);
}

// ------------------------------------------------------------------------
// Panel LU kernels with the rows left in place (k_lu_panel_rr for panels of
// at most 512 rows, k_lu_panel_ll above).  One workgroup per front.  Slot
// s = q * PNT + tid is panel row s (front row c0 + s) and belongs to thread
// tid.  Rows never move while the panel is factored: each thread keeps the
// current position of its slots (the swap g <-> p of step g acts on
// positions) and the step at which a slot became a pivot, and the rows are
// stored at their final positions once, at the end.  The columns run in
// blocks of W; a column step is a wavefront argmax (DPP), one barrier that
// exchanges the wavefronts' candidates (double-buffered in LDS), and the
// rank-1 update of the block's later columns in registers.  The kb pivot
// rows live in LDS in pivot order (s_pr, all kb columns); after a block,
// its pivot rows receive the updates of the block's own pivots for the
// later columns (U strip).  Every entry sees the rank-1 update of each
// earlier pivot as one FMA, in pivot order: the operation sequence of the
// unblocked panel (getf2; FLA_LU_piv on the panel,
// LINAL/src/dense/lu_piv.cxx:79) and of k_lu_panel, so FP64 fronts get the
// same bits from all three kernels.  Pivot rule: first maximal |a| in
// current row order; a column without a comparable entry (all NaN) keeps
// the current row.
// ------------------------------------------------------------------------
template <int N, typename Fn, int... I>
__device__ __forceinline__ void static_for_impl(Fn&& fn, std::integer_sequence<int, I...>) {
  (fn(std::integral_constant<int, I>{}), ...);
}
// fn(integral_constant<int, i>) for i = 0 .. N-1, every index a compile-time constant
template <int N, typename Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
  static_for_impl<N>(fn, std::make_integer_sequence<int, N>{});
}

constexpr int PL_BIG = 0x3fffffff;  // step of a slot that is not (yet) a pivot
constexpr int PL_MAXROWS = 8192;    // pivot-candidate rows of the panel kernels
constexpr int PL_LDR = TILE + 1;    // row stride of the pivot rows in LDS

// Workgroup barrier for LDS traffic only: LDS writes complete before it,
// global loads in flight stay in flight (__syncthreads would wait for them).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// x combined with the value DPP control CTRL moves into each lane (lanes
// outside ROW_MASK keep their own value)
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_max_f64(double x) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(x), __double2loint(x), CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(x), __double2hiint(x), CTRL, ROW_MASK, 0xF, false);
  return fmax(x, __hiloint2double(hi, lo));
}

// Largest value over the wavefront (no NaN among the inputs): quad swaps,
// half-row and row mirrors, then the row broadcasts 15 and 31 carry the
// maximum to lane 63.
__device__ __forceinline__ double wave_max_f64(double x) {
  x = dpp_max_f64<0xB1>(x);        // quad_perm [1, 0, 3, 2]
  x = dpp_max_f64<0x4E>(x);        // quad_perm [2, 3, 0, 1]
  x = dpp_max_f64<0x141>(x);       // row_half_mirror
  x = dpp_max_f64<0x140>(x);       // row_mirror
  x = dpp_max_f64<0x142, 0xA>(x);  // row_bcast:15 into rows 1, 3
  x = dpp_max_f64<0x143, 0xC>(x);  // row_bcast:31 into rows 2, 3
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), 63),
                          __builtin_amdgcn_readlane(__double2loint(x), 63));
}

// Wavefront argmax of (value, position, slot): the largest value, the
// smallest position among equal values.  cv is never NaN (candidates start
// at -1 and only take comparable values).
__device__ __forceinline__ void wave_argmax(double& cv, int& cp, int& cs) {
  const double mx = wave_max_f64(cv);
  const unsigned long long tie = __ballot(cv == mx);
  if ((tie & (tie - 1ull)) == 0ull) {
    const int l = __ffsll((unsigned long long)tie) - 1;
    cp = __builtin_amdgcn_readlane(cp, l);
    cs = __builtin_amdgcn_readlane(cs, l);
  } else {
    int p = (cv == mx) ? cp : 0x7fffffff;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) p = min(p, __shfl_xor(p, o));
    const unsigned long long w = __ballot(cv == mx && cp == p);
    const int l = __ffsll((unsigned long long)w) - 1;
    cs = __builtin_amdgcn_readlane(cs, l);
    cp = p;
  }
  cv = mx;
}

struct alignas(16) PanelCand {
  double v;  // largest |a| of the wavefront's rows (-1: none)
  int p, s;  // its position and slot
};

template <int W, int NW>
struct PanelLds {
  double pr[TILE * PL_LDR];  // pivot rows in pivot order
  PanelCand cand[2][NW];
  double crow[2][NW + 1][W];
  int piv[TILE], ps[TILE];  // relative pivot and pivot slot of each step
  int dsl[TILE], dpos[TILE], tgt[TILE];
  int nd, ncnt, keep;
};

// Column step K of the block starting at column j0 (step g = j0 + K): the
// pivot among the slots not yet pivoted, the bookkeeping, and the rank-1
// update of the block's later columns of every slot not yet pivoted.
// v[q][j] is column j0 + j of slot q * PNT + tid; pst[q] is the step at
// which that slot became a pivot (PL_BIG: not yet, -1: no row).  Only
// slots below kb <= 64 (q = 0 of the first wavefront) can ever be
// displaced, so a slot's position is its index except for q = 0, whose
// position is pos0.  Returns the pivot slot; pr[j] (j >= K) receives the
// pivot row's block values, and so does L.pr (all W of them).
template <int K, int R, int W, int PNT>
__device__ __forceinline__ int panel_step(double (&v)[R][W], int& pos0, int (&pst)[R], double (&pr)[W], int j0,
                                          PanelLds<W, PNT / 64>& L, int tid, int& zero) {
  constexpr int NW = PNT / 64;
  const int g = j0 + K, buf = g & 1, lane = tid & 63, wid = tid >> 6;
  // this thread's candidate: first maximal |a| in position order (slots
  // q >= 1 sit at their index, which grows with q)
  double cv = -1.0;
  int cs = -1;
#pragma unroll
  for (int q = 1; q < R; ++q) {
    const double a = fabs(v[q][K]);
    if (pst[q] == PL_BIG && a > cv) {
      cv = a;
      cs = q * PNT + tid;
    }
  }
  int cp = cs >= 0 ? cs : 0x7fffffff;
  {
    const double a = fabs(v[0][K]);
    if (pst[0] == PL_BIG && (a > cv || (a == cv && pos0 < cp))) {
      cv = a;
      cp = pos0;
      cs = tid;
    }
  }
  wave_argmax(cv, cp, cs);
  if constexpr (NW == 1 && R == 1) {
    // one wavefront: its argmax is the pivot; the row comes from the
    // owner's lane without an LDS round trip
    int bs = cs, bp = cp;
    if (bs < 0) {
      // no comparable entry in the column: the row at position g stays
      bs = __ffsll((unsigned long long)__ballot(pst[0] == PL_BIG && pos0 == g)) - 1;
      bp = g;
    }
#pragma unroll
    for (int j = K; j < W; ++j)
      pr[j] = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v[0][j]), bs),
                               __builtin_amdgcn_readlane(__double2loint(v[0][j]), bs));
    if (tid == bs)
#pragma unroll
      for (int j = 0; j < W; ++j) L.pr[g * PL_LDR + j0 + j] = v[0][j];
    const double pv = pr[K], rinv = 1.0 / pv;
    if (!(pv != 0.0) && zero == 0) zero = g + 1;
    if (tid == 0) {
      L.piv[g] = bp - g;
      L.ps[g] = bs;
    }
    if (tid == bs) pst[0] = g;
    if (pst[0] == PL_BIG && pos0 == g) pos0 = bp;  // displaced by the pivot row
    if (pst[0] == PL_BIG) {
      const double l = lu_mult(v[0][K], pv, rinv);
      v[0][K] = l;
#pragma unroll
      for (int j = K + 1; j < W; ++j) v[0][j] = fma(-l, pr[j], v[0][j]);
    }
    return bs;
  }
  if (lane == 0) L.cand[buf][wid] = PanelCand{cv, cp, cs};
  // the wavefront's best row, published by its owner (a uniform register
  // index here became a dynamic index into scratch memory)
#pragma unroll
  for (int q = 0; q < R; ++q)
    if (q * PNT + tid == cs)
#pragma unroll
      for (int j = 0; j < W; ++j) L.crow[buf][wid][j] = v[q][j];
  lds_barrier();
  // the best of the NW candidates: lane w < NW holds candidate w
  int bs, bp, bw;
  {
    const PanelCand c = (lane < NW) ? L.cand[buf][lane < NW ? lane : 0] : PanelCand{-1.0, 0x7fffffff, -1};
    double mv = c.v;
    if constexpr (NW > 1) mv = dpp_max_f64<0xB1>(mv);   // quad_perm [1, 0, 3, 2]
    if constexpr (NW > 2) mv = dpp_max_f64<0x4E>(mv);   // quad_perm [2, 3, 0, 1]
    if constexpr (NW > 4) mv = dpp_max_f64<0x141>(mv);  // row_half_mirror
    mv = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(mv), 0),
                          __builtin_amdgcn_readlane(__double2loint(mv), 0));
    const bool top = lane < NW && c.v == mv;
    const unsigned long long tie = __ballot(top);
    if ((tie & (tie - 1ull)) == 0ull) {
      bw = __ffsll((unsigned long long)tie) - 1;
    } else {
      int pm = top ? c.p : 0x7fffffff;
#pragma unroll
      for (int o = 4; o > 0; o >>= 1) pm = min(pm, __shfl_xor(pm, o));
      pm = __builtin_amdgcn_readlane(pm, 0);
      bw = __ffsll((unsigned long long)__ballot(top && c.p == pm)) - 1;
    }
    bp = __builtin_amdgcn_readlane(c.p, bw);
    bs = __builtin_amdgcn_readlane(c.s, bw);
  }
  if (bs < 0) {
    // no comparable entry in the column: the row at position g stays, and
    // its owner publishes it
    if (tid < 64 && pst[0] == PL_BIG && pos0 == g) {
      L.keep = tid;
#pragma unroll
      for (int j = 0; j < W; ++j) L.crow[buf][NW][j] = v[0][j];
    }
    lds_barrier();
    bs = L.keep;
    bp = g;
    bw = NW;
  }
#pragma unroll
  for (int j = K; j < W; ++j) pr[j] = L.crow[buf][bw][j];
  const double pv = pr[K], rinv = 1.0 / pv;
  if (!(pv != 0.0) && zero == 0) zero = g + 1;
  // the pivot row's block values (L part left of K, U part from K on)
  if (tid < W) L.pr[g * PL_LDR + j0 + tid] = L.crow[buf][bw][tid];
  if (tid == 0) {
    L.piv[g] = bp - g;
    L.ps[g] = bs;
  }
#pragma unroll
  for (int q = 0; q < R; ++q)
    if (q * PNT + tid == bs) pst[q] = g;
  if (tid < 64 && pst[0] == PL_BIG && pos0 == g) pos0 = bp;  // displaced by the pivot row
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (pst[q] == PL_BIG) {
      const double l = lu_mult(v[q][K], pv, rinv);
      v[q][K] = l;
#pragma unroll
      for (int j = K + 1; j < W; ++j) v[q][j] = fma(-l, pr[j], v[q][j]);
    }
  }
  return bs;
}

// U strip of the block's pivot rows [j0, j0 + nbr) for the columns
// [j1, kb): the rows in L.pr hold the updates of every pivot before j0;
// the block's own pivots are applied here, in order, one thread per column.
template <int W, int PNT>
__device__ __forceinline__ void panel_ustrip(PanelLds<W, PNT / 64>& L, int j0, int nbr, int j1, int kb, int tid) {
  for (int c = j1 + tid; c < kb; c += PNT) {
    double u[W];
#pragma unroll
    for (int r = 0; r < W; ++r) u[r] = (r < nbr) ? L.pr[(j0 + r) * PL_LDR + c] : 0.0;
#pragma unroll
    for (int r = 1; r < W; ++r) {
      // the block's L row r, read just before its use (hoisting every row's
      // reads at once costs more registers than the kernel has)
      __builtin_amdgcn_sched_barrier(0);
      const double* __restrict__ lr = L.pr + (j0 + r) * PL_LDR + j0;
#pragma unroll
      for (int i = 0; i < r; ++i) u[r] = fma(-lr[i], u[i], u[r]);
    }
#pragma unroll
    for (int r = 0; r < W; ++r)
      if (r < nbr) L.pr[(j0 + r) * PL_LDR + c] = u[r];
  }
}

// Relative pivots, the first zero pivot and the net row permutation for the
// row kernel (the layout of panel_permutation): the rows whose content
// changed (positions [0, kb), then the distinct swap targets beyond kb in
// order of first use) and, per listed row, the list index of the row whose
// old content lands there.  The final slot at position g < kb is the
// pivot of step g (L.ps); L.dsl / L.dpos hold the displaced slots (not
// pivots, moved; all below kb) and their final positions, which are the
// swap targets beyond kb.  A pivot slot at or beyond kb was never moved
// before its step, so it is a swap target too.  mapS / mapU (LDS, m bytes
// each, no initialization needed: only target positions are read) take
// target position -> displaced slot and target position -> list index.
template <int W, int PNT>
__device__ __forceinline__ void panel_finish(PanelLds<W, PNT / 64>& L, const uhmk_front& fr, int f, int c0, int kb,
                                             int zero, int* __restrict__ piv, int* __restrict__ info,
                                             int* __restrict__ pperm, unsigned char* mapS, unsigned char* mapU,
                                             int tid) {
  for (int k = tid; k < kb; k += PNT) piv[fr.poff + c0 + k] = L.piv[k];
  if (tid == 0 && zero && info[f] == 0) info[f] = c0 + zero;
  lds_barrier();  // the maps may overlay LDS that was read until here
  if (tid < 64) {
    const int k = tid;
    const int t = (k < kb) ? k + L.piv[k] : -1;
    bool first = t >= kb;
#pragma unroll
    for (int k2 = 0; k2 < 63; ++k2) first = first && !(k2 < k && __builtin_amdgcn_readlane(t, k2) == t);
    const unsigned long long mask = __ballot(first);
    if (first) {
      const int u = __popcll(mask & ((1ull << k) - 1ull));
      L.tgt[u] = t;
      mapU[t] = (unsigned char)u;
    }
    if (k == 0) L.ncnt = kb + __popcll(mask);
  }
  const int nd = min(L.nd, TILE);
  for (int d = tid; d < nd; d += PNT) mapS[L.dpos[d]] = (unsigned char)L.dsl[d];
  lds_barrier();
  const int cnt = L.ncnt;
  int* __restrict__ pp = pperm + (long long)f * PP_STRIDE;
  for (int k = tid; k < cnt; k += PNT) {
    const int lr = (k < kb) ? k : L.tgt[k - kb];
    const int src = (lr < kb) ? L.ps[lr] : (int)mapS[lr];
    const int si = (src < kb) ? src : kb + (int)mapU[src];
    pp[1 + k] = c0 + lr;
    pp[1 + 2 * TILE + k] = si;
  }
  if (tid == 0) pp[0] = cnt;
}

// ---- panels of at most PNT rows: every row's kb columns in registers.
// Right-looking across the blocks: at the end of a block the pivot rows
// publish their later columns, the U strip applies the block's own pivots
// to them, and every slot not yet pivoted applies the block's rank-W update
// to its later columns (L in its registers, U from LDS).  No memory traffic
// between the first load and the final stores.
template <typename T, int PNT>
__global__ __launch_bounds__(PNT) void k_lu_panel_rr(const uhmk_front* __restrict__ F,
                                                     const int* __restrict__ flist, int c0, int nb,
                                                     T* __restrict__ arena, int* __restrict__ piv,
                                                     int* __restrict__ info, int* __restrict__ pperm) {
  constexpr int W = 16, NB = TILE / W, NW = PNT / 64;
  __shared__ PanelLds<W, NW> L;
  __shared__ unsigned char s_map[2 * PNT];
  const int f = flist[blockIdx.x];
  const uhmk_front fr = F[f];
  const int n = fr.n, fs = fr.fs;
  const int m = fs - c0, kb = min(nb, m);
  T* __restrict__ P = arena + fr.off + c0 + (long long)c0 * n;  // panel origin (c0, c0)
  const int tid = threadIdx.x;
  int pos0 = tid, pst[1] = {tid < m ? PL_BIG : -1};
  if (tid == 0) L.nd = 0;
  double a[TILE];
#pragma unroll
  for (int j = 0; j < TILE; ++j) a[j] = uhmd::ldz(P, tid + (long long)j * n, tid < m && j < kb);
  int zero = 0;
  static_for<NB>([&](auto bc) {
    constexpr int B = decltype(bc)::value, J0 = B * W, J1 = J0 + W;
    if (J0 >= kb) return;
    double v[1][W], pr[W];
#pragma unroll
    for (int j = 0; j < W; ++j) v[0][j] = a[J0 + j];
    const int kend = min(W, kb - J0);
    static_for<W>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      if (K >= kend) return;
      panel_step<K, 1, W, PNT>(v, pos0, pst, pr, J0, L, tid, zero);
    });
#pragma unroll
    for (int j = 0; j < W; ++j) a[J0 + j] = v[0][j];
    if (J1 < kb) {
      // the block's pivot rows publish their later columns
      if (pst[0] >= J0 && pst[0] < J1)
#pragma unroll
        for (int c = J1; c < TILE; ++c) L.pr[pst[0] * PL_LDR + c] = a[c];
      lds_barrier();
      panel_ustrip<W, PNT>(L, J0, W, J1, kb, tid);
      lds_barrier();
      if (pst[0] == PL_BIG) {
#pragma unroll
        for (int i = 0; i < W; ++i) {
          const double li = a[J0 + i];
          const double* __restrict__ ur = L.pr + (J0 + i) * PL_LDR;
#pragma unroll
          for (int c = J1; c < TILE; ++c) a[c] = fma(-li, ur[c], a[c]);
        }
      }
    }
  });
  __syncthreads();
  // final placement: pivot rows (registers for the columns up to the end of
  // their block, LDS beyond) at their step, the other rows at their position
  if (pst[0] >= 0) {
    const int g = pst[0] == PL_BIG ? pos0 : pst[0];
    const int j1 = pst[0] == PL_BIG ? TILE : (pst[0] / W + 1) * W;
#pragma unroll
    for (int c = 0; c < TILE; ++c)
      if (c < kb) P[g + (long long)c * n] = (T)(c < j1 ? a[c] : L.pr[g * PL_LDR + c]);
    if (pst[0] == PL_BIG && pos0 != tid) {
      const int d = atomicAdd(&L.nd, 1);
      if (d < TILE) {
        L.dsl[d] = tid;
        L.dpos[d] = pos0;
      }
    }
  }
  __syncthreads();
  panel_finish<W, PNT>(L, fr, f, c0, kb, zero, piv, info, pperm, s_map, s_map + PNT, tid);
}

// v -= L(slot, [i0, i0 + LC)) U([i0, i0 + LC), block) for every slot, one
// U row at a time (reading every row's values ahead of the FMAs costs more
// registers than the kernels have); rows of pivoted slots hold zeros in l.
template <int R, int W, int LC, int PNT>
__device__ __forceinline__ void panel_deferred(double (&v)[R][W], const double (&l)[R][LC],
                                               const PanelLds<W, PNT / 64>& L, int i0, int j0) {
#pragma unroll
  for (int u = 0; u < LC; ++u) {
    __builtin_amdgcn_sched_barrier(0);
    const double* __restrict__ ur = L.pr + (i0 + u) * PL_LDR + j0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const double uu = ur[j];
#pragma unroll
      for (int q = 0; q < R; ++q) v[q][j] = fma(-l[q][u], uu, v[q][j]);
    }
  }
}

// ---- taller panels: left-looking over register blocks of W columns (R x W
// per thread).  A block first receives the updates of every earlier pivot
// in pivot order (L of the slot from memory, where the slot's earlier
// blocks were stored, LC columns at a time; U rows from LDS), then its
// column steps run.  The block's pivot rows are
// fetched for the other columns while the later steps run (their L part
// and their raw values right of the block), then receive the updates of
// the earlier pivots (U strip, two phases).
template <int R, int W, int LC, bool PF, typename T, int PNT>
__global__ __launch_bounds__(PNT) void k_lu_panel_ll(const uhmk_front* __restrict__ F,
                                                     const int* __restrict__ flist, int c0, int nb,
                                                     T* __restrict__ arena, int* __restrict__ piv,
                                                     int* __restrict__ info, int* __restrict__ pperm) {
  constexpr int NW = PNT / 64;
  __shared__ PanelLds<W, NW> L;
  __shared__ double s_stage[TILE * PL_LDR];  // displaced rows for the final placement (padded rows)

  const int f = flist[blockIdx.x];
  const uhmk_front fr = F[f];
  const int n = fr.n, fs = fr.fs;
  const int m = fs - c0, kb = min(nb, m);
  T* __restrict__ P = arena + fr.off + c0 + (long long)c0 * n;  // panel origin (c0, c0)
  const int tid = threadIdx.x;

  int pos0 = tid, pst[R];
#pragma unroll
  for (int q = 0; q < R; ++q) pst[q] = q * PNT + tid < m ? PL_BIG : -1;
  if (tid == 0) L.nd = 0;
  int zero = 0;
  double v[R][W];
  for (int j0 = 0; j0 < kb; j0 += W) {
    const int j1 = min(j0 + W, kb);
    const int nbr = j1 - j0, wl = j0, wr = kb - j1;
    // raw values of the block for the slots not yet pivoted
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int s = q * PNT + tid;
      const bool ok = pst[q] == PL_BIG;
#pragma unroll
      for (int j = 0; j < W; ++j) v[q][j] = uhmd::ldz(P, s + (long long)(j0 + j) * n, ok && j0 + j < kb);
    }
    // updates of the pivots [0, j0), in pivot order, LC L columns at a
    // time; with PF the next LC columns are loaded ahead of the FMAs (the
    // taller shapes have no registers left for that)
    if constexpr (PF) {
      if (j0 > 0) {
        double l[R][LC], ln[R][LC];
#pragma unroll
        for (int q = 0; q < R; ++q)
#pragma unroll
          for (int u = 0; u < LC; ++u) l[q][u] = uhmd::ldz(P, q * PNT + tid + (long long)u * n, pst[q] == PL_BIG);
        for (int i0 = 0; i0 < j0; i0 += LC) {
          const bool more = i0 + LC < j0;
#pragma unroll
          for (int q = 0; q < R; ++q)
#pragma unroll
            for (int u = 0; u < LC; ++u)
              ln[q][u] = uhmd::ldz(P, q * PNT + tid + (long long)(i0 + LC + u) * n, more && pst[q] == PL_BIG);
          panel_deferred<R, W, LC, PNT>(v, l, L, i0, j0);
#pragma unroll
          for (int q = 0; q < R; ++q)
#pragma unroll
            for (int u = 0; u < LC; ++u) l[q][u] = ln[q][u];
        }
      }
    } else {
      for (int i0 = 0; i0 < j0; i0 += LC) {
        double l[R][LC];
#pragma unroll
        for (int q = 0; q < R; ++q)
#pragma unroll
          for (int u = 0; u < LC; ++u)
            l[q][u] = uhmd::ldz(P, q * PNT + tid + (long long)(i0 + u) * n, pst[q] == PL_BIG);
        panel_deferred<R, W, LC, PNT>(v, l, L, i0, j0);
      }
    }
    // the block's column steps; after step K, thread t < wl + wr starts the
    // load of the pivot row's column col(t) (its L part left of the block,
    // its raw values right of it)
    const int kend = min(W, kb - j0);
    const int myc = tid < wl ? tid : j1 + (tid - wl);
    const bool fetch = tid < wl + wr;
    double pa[W], pr[W];
    static_for<W>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      if (K >= kend) return;
      const int bs = panel_step<K, R, W, PNT>(v, pos0, pst, pr, j0, L, tid, zero);
      pa[K] = fetch ? (double)P[bs + (long long)myc * n] : 0.0;
    });
    // store the block for the slots not pivoted before it
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int s = q * PNT + tid;
      if (pst[q] >= j0)
#pragma unroll
        for (int j = 0; j < W; ++j)
          if (j0 + j < kb) P[s + (long long)(j0 + j) * n] = (T)v[q][j];
    }
    if (fetch)
#pragma unroll
      for (int k = 0; k < W; ++k)
        if (k < nbr) L.pr[(j0 + k) * PL_LDR + myc] = pa[k];
    // the block's stores reach memory before any thread reads them back
    // (the block's pivot rows, the deferred updates of the next blocks)
    __syncthreads();
    if (wr > 0) {
      // U strip, phase 1: the pivots before the block, one thread per entry
      if (j0 > 0) {
        for (int idx = tid; idx < nbr * wr; idx += PNT) {
          const int g = j0 + idx % nbr, c = j1 + idx / nbr;
          double acc = L.pr[g * PL_LDR + c];
          const double* __restrict__ lr = L.pr + g * PL_LDR;
#pragma unroll 8
          for (int i = 0; i < j0; ++i) acc = fma(-lr[i], L.pr[i * PL_LDR + c], acc);
          L.pr[g * PL_LDR + c] = acc;
        }
        lds_barrier();
      }
      // phase 2: the block's own pivots
      panel_ustrip<W, PNT>(L, j0, nbr, j1, kb, tid);
      lds_barrier();
    }
  }
  __syncthreads();
  // final placement: pivot rows to positions [0, kb), displaced rows (not
  // pivots, position changed) staged through LDS to their positions
  if (pst[0] == PL_BIG && pos0 != tid) {
    // at most kb: only slots that held a position in [0, kb) get displaced
    const int d = atomicAdd(&L.nd, 1);
    if (d < TILE) {
      L.dsl[d] = tid;
      L.dpos[d] = pos0;
    }
  }
  __syncthreads();
  const int nd = min(L.nd, TILE);
  for (int idx = tid; idx < nd * kb; idx += PNT) {
    const int d = idx % nd, c = idx / nd;
    s_stage[d * PL_LDR + c] = (double)P[L.dsl[d] + (long long)c * n];
  }
  __syncthreads();
  for (int idx = tid; idx < kb * kb; idx += PNT) {
    const int g = idx % kb, c = idx / kb;
    P[g + (long long)c * n] = (T)L.pr[g * PL_LDR + c];
  }
  for (int idx = tid; idx < nd * kb; idx += PNT) {
    const int d = idx % nd, c = idx / nd;
    P[L.dpos[d] + (long long)c * n] = (T)s_stage[d * PL_LDR + c];
  }
  unsigned char* map = reinterpret_cast<unsigned char*>(s_stage);
  panel_finish<W, PNT>(L, fr, f, c0, kb, zero, piv, info, pperm, map, map + PL_MAXROWS, tid);
}


template <typename T>
int launch_rows(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, T* arena,
                const int* pperm, int width, hipStream_t stream) {
  if (width == 16)
    hipLaunchKernelGGL((k_lu_rows<T, 64>), dim3(n_items), dim3(64), 0, stream, fronts, items, c0, nb, arena, pperm);
  else if (width == 32)
    hipLaunchKernelGGL((k_lu_rows<T, 128>), dim3(n_items), dim3(128), 0, stream, fronts, items, c0, nb, arena, pperm);
  else if (width == 64)
    hipLaunchKernelGGL((k_lu_rows<T, 256>), dim3(n_items), dim3(256), 0, stream, fronts, items, c0, nb, arena, pperm);
  else
    return -1000;
  return ret(hipGetLastError());
}

template <typename T>
int launch_panel_ll(const uhmk_front* fronts, const int* flist, int n_fronts, int max_rows, int c0, int nb,
                    T* arena, int* piv, int* info, int* pperm, hipStream_t stream) {
  const dim3 g(n_fronts);
#define UHM_PANEL_RR(PNT)                                                                                        \
  hipLaunchKernelGGL((k_lu_panel_rr<T, PNT>), g, dim3(PNT), 0, stream, fronts, flist, c0, nb, arena, piv, info, \
                     pperm)
#define UHM_PANEL_LL(R, W, LC, PF, PNT)                                                                \
  hipLaunchKernelGGL((k_lu_panel_ll<R, W, LC, PF, T, PNT>), g, dim3(PNT), 0, stream, fronts, flist, c0, nb, arena, \
                     piv, info, pperm)
  if (max_rows <= 64) UHM_PANEL_RR(64);
  else if (max_rows <= 128) UHM_PANEL_RR(128);
  else if (max_rows <= 256) UHM_PANEL_RR(256);
  else if (max_rows <= 512) UHM_PANEL_RR(512);
  else if (max_rows <= 1024) UHM_PANEL_LL(2, 8, 8, true, 512);
  else if (max_rows <= 1536) UHM_PANEL_LL(3, 8, 8, true, 512);
  else if (max_rows <= 2048) UHM_PANEL_LL(4, 8, 4, true, 512);
  else if (max_rows <= 3072) UHM_PANEL_LL(6, 8, 4, false, 512);
  else if (max_rows <= 4096) UHM_PANEL_LL(8, 8, 1, false, 512);
  else if (max_rows <= 6144) UHM_PANEL_LL(12, 4, 2, false, 512);
  else if (max_rows <= PL_MAXROWS) UHM_PANEL_LL(16, 2, 1, false, 512);
  else return -1001;  // more pivot-candidate rows than the panel kernels hold
#undef UHM_PANEL_LL
#undef UHM_PANEL_RR
  return ret(hipGetLastError());
}

template <typename T>
int launch_panel(const uhmk_front* fronts, const int* flist, int n_fronts, int max_rows, int c0, int nb, T* arena,
                 int* piv, int* info, int* pperm, hipStream_t stream) {
  dim3 g(n_fronts), b(512);
  if (max_rows <= 64)
    hipLaunchKernelGGL((k_lu_panel_reg<T, 64>), g, dim3(64), 0, stream, fronts, flist, c0, nb, arena, piv, info,
                       pperm);
  else if (max_rows <= 128)
    hipLaunchKernelGGL((k_lu_panel_reg<T, 128>), g, dim3(128), 0, stream, fronts, flist, c0, nb, arena, piv, info,
                       pperm);
  else if (max_rows <= 256)
    hipLaunchKernelGGL((k_lu_panel_reg<T, 256>), g, dim3(256), 0, stream, fronts, flist, c0, nb, arena, piv, info,
                       pperm);
  else if (max_rows <= 512)
    hipLaunchKernelGGL((k_lu_panel_reg<T, 512>), g, dim3(512), 0, stream, fronts, flist, c0, nb, arena, piv, info,
                       pperm);
  else if (max_rows <= 1024)
    hipLaunchKernelGGL((k_lu_panel<2, T, 512>), g, b, 0, stream, fronts, flist, c0, nb, arena, piv, info, pperm);
  else if (max_rows <= 2048)
    hipLaunchKernelGGL((k_lu_panel<4, T, 512>), g, b, 0, stream, fronts, flist, c0, nb, arena, piv, info, pperm);
  else if (max_rows <= 3072)
    hipLaunchKernelGGL((k_lu_panel<6, T, 512>), g, b, 0, stream, fronts, flist, c0, nb, arena, piv, info, pperm);
  else if (max_rows <= 4096)
    hipLaunchKernelGGL((k_lu_panel<8, T, 512>), g, b, 0, stream, fronts, flist, c0, nb, arena, piv, info, pperm);
  else if (max_rows <= 6144)
    hipLaunchKernelGGL((k_lu_panel<12, T, 512>), g, b, 0, stream, fronts, flist, c0, nb, arena, piv, info, pperm);
  else
    return -1001;  // more than 6144 pivot-candidate rows
  return ret(hipGetLastError());
}

// ------------------------------------------------------------------------
// Whole-front LU for fronts with n <= 64 (the leaves and other small
// fronts): one wavefront per front, lane i holds row i of the assembled
// front (all n columns) in registers, and the fs pivot steps run
// right-looking over the whole front, so that one launch produces L\U in
// ATL, U12 in ATR, L21 in ABL and the Schur complement in ABR
// (lu_piv_hier, UHM/src/matrix/uhm/fla/lu_piv/decompose.cxx:124-153, for a
// front that fits one panel).  Rows are never moved: every lane tracks its
// row's current position (the swaps k <-> p of the pivot sequence act on
// positions) and stores the row at its final position at the end.  The
// pivot rule and arithmetic are those of k_lu_panel / k_lu_rows: pivots
// searched in the ATL rows only, first maximal |a| in current order, ATL
// multipliers as lu_mult forms them (LAPACK dgetf2), ABL multipliers by the
// reciprocal of the pivot, every update a rank-1 FMA in pivot order.
// ------------------------------------------------------------------------
__device__ __forceinline__ double lane_d(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

template <typename T>
__global__ __launch_bounds__(256) void k_lu_front64(const uhmk_front* __restrict__ F,
                                                    const int* __restrict__ flist, int n_fronts,
                                                    T* __restrict__ arena, int* __restrict__ piv,
                                                    int* __restrict__ info) {
  const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= n_fronts) return;
  const int f = flist[q];
  const uhmk_front fr = F[f];
  const int n = fr.n, fs = fr.fs;
  const int lane = threadIdx.x & 63;
  T* __restrict__ A = arena + fr.off;
  double a[TILE];
#pragma unroll
  for (int j = 0; j < TILE; ++j) a[j] = uhmd::ldz(A, lane + (long long)j * n, lane < n && j < n);
  int pos = lane;
  int zero = 0;
  const bool abl = lane >= fs;
  int* __restrict__ p = piv + fr.poff;
  // columns in groups of 8: group index at run time, column inside the
  // group and every register index at compile time
  for (int g = 0; 8 * g < fs; ++g) {
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      const int K = 8 * g + kk;
      if (K >= fs) break;
      double ak = 0.0;
#pragma unroll
      for (int h = 0; h < 8; ++h)
        if (h == g) ak = a[8 * h + kk];
      // first maximal |a(., K)| over the rows at positions [K, fs)
      double cv = -1.0;
      int ck = 0x7fffffff;
      if (pos >= K && pos < fs) {
        const double v = fabs(ak);
        if (v > cv) {
          cv = v;
          ck = pos;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(cv, o);
        const int ok = __shfl_xor(ck, o);
        if (ov > cv || (ov == cv && ok < ck)) {
          cv = ov;
          ck = ok;
        }
      }
      // a column of NaNs has no maximum: keep the current row
      const int pp = (ck < K || ck >= fs) ? K : ck;
      const int P = __ffsll((long long)__ballot(pos == pp)) - 1;
      const int C = __ffsll((long long)__ballot(pos == K)) - 1;
      const double pv = lane_d(ak, P);
      if (lane == 0) {
        p[K] = pp - K;
        if (!(pv != 0.0) && zero == 0) zero = K + 1;
      }
      if (P != C) {
        if (lane == P) pos = K;
        else if (lane == C) pos = pp;
      }
      const bool upd = pos > K;
      double l = 0.0;
      if (upd) {
        l = ak;
        if (abl) l = l * (1.0 / pv);
        else l = lu_mult(l, pv, 1.0 / pv);
      }
#pragma unroll
      for (int h = 0; h < 8; ++h) {
        if (h < g) continue;
        if (h == g && upd) a[8 * h + kk] = l;
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const int j = 8 * h + jj;
          if ((h > g || jj > kk) && j < n) {
            const double u = lane_d(a[j], P);
            if (upd) a[j] = fma(-l, u, a[j]);
          }
        }
      }
    }
  }
  if (lane < n)
#pragma unroll
    for (int j = 0; j < TILE; ++j)
      if (j < n) A[pos + (long long)j * n] = (T)a[j];
  if (lane == 0 && zero && info[f] == 0) info[f] = zero;
}

}  // namespace

extern "C" {

int uhmk_copy_in(const uhmk_front* fronts, const uhmk_item* items, int n_items, const uhmk_cblock* blocks,
                 const double* vals, const int* maps, double* stage, double* b, int nrhs, hipStream_t stream) {
  if (n_items <= 0) return 0;
  hipLaunchKernelGGL(k_copy_in, dim3(n_items), dim3(256), 0, stream, fronts, items, blocks, vals, maps, stage, b, nrhs);
  return ret(hipGetLastError());
}

int uhmk_copy_front(const void* src, int src_prec, void* dst, int dst_prec, long long count, hipStream_t stream) {
  if (count <= 0) return 0;
  if (src_prec < 0 || src_prec > 1 || dst_prec < 0 || dst_prec > 1) return -1006;
  const dim3 g((unsigned)((count + 255) / 256)), b(256);
  if (src_prec == 0 && dst_prec == 0)
    hipLaunchKernelGGL((k_copy_cast<double, double>), g, b, 0, stream, static_cast<const double*>(src),
                       static_cast<double*>(dst), count);
  else if (src_prec == 0)
    hipLaunchKernelGGL((k_copy_cast<double, float>), g, b, 0, stream, static_cast<const double*>(src),
                       static_cast<float*>(dst), count);
  else if (dst_prec == 0)
    hipLaunchKernelGGL((k_copy_cast<float, double>), g, b, 0, stream, static_cast<const float*>(src),
                       static_cast<double*>(dst), count);
  else
    hipLaunchKernelGGL((k_copy_cast<float, float>), g, b, 0, stream, static_cast<const float*>(src),
                       static_cast<float*>(dst), count);
  return ret(hipGetLastError());
}

int uhmk_assemble(const uhmk_front* fronts, const uhmk_item* items, int n_items, const int* src_beg,
                  const uhmk_source* srcs, const int* inv, double* arena64, float* arena32, int prec,
                  const double* stage, hipStream_t stream) {
  if (n_items <= 0) return 0;
  if (prec == 0)
    hipLaunchKernelGGL(k_assemble<double>, dim3(n_items), dim3(256), 0, stream, fronts, items, src_beg, srcs, inv,
                       arena64, arena32, stage);
  else if (prec == 1)
    hipLaunchKernelGGL(k_assemble<float>, dim3(n_items), dim3(256), 0, stream, fronts, items, src_beg, srcs, inv,
                       arena64, arena32, stage);
  else
    return -1006;
  return ret(hipGetLastError());
}

int uhmk_lu_panel(const uhmk_front* fronts, const int* flist, int n_fronts, int max_rows, int c0, int nb,
                  void* arena, int prec, int* piv, int* info, int* pperm, hipStream_t stream) {
  if (n_fronts <= 0) return 0;
  if (nb > TILE || nb < 1) return -1000;
  if (prec == 0)
    return launch_panel<double>(fronts, flist, n_fronts, max_rows, c0, nb, static_cast<double*>(arena), piv, info,
                                pperm, stream);
  if (prec == 1)
    return launch_panel<float>(fronts, flist, n_fronts, max_rows, c0, nb, static_cast<float*>(arena), piv, info,
                               pperm, stream);
  return -1006;
}

int uhmk_lu_panel_ll(const uhmk_front* fronts, const int* flist, int n_fronts, int max_rows, int c0, int nb,
                     void* arena, int prec, int* piv, int* info, int* pperm, hipStream_t stream) {
  if (n_fronts <= 0) return 0;
  if (nb > TILE || nb < 1) return -1000;
  if (prec == 0)
    return launch_panel_ll<double>(fronts, flist, n_fronts, max_rows, c0, nb, static_cast<double*>(arena), piv, info,
                                   pperm, stream);
  if (prec == 1)
    return launch_panel_ll<float>(fronts, flist, n_fronts, max_rows, c0, nb, static_cast<float*>(arena), piv, info,
                                  pperm, stream);
  return -1006;
}

int uhmk_lu_front64(const uhmk_front* fronts, const int* flist, int n_fronts, void* arena, int prec, int* piv,
                    int* info, hipStream_t stream) {
  if (n_fronts <= 0) return 0;
  const dim3 g((n_fronts + 3) / 4), b(256);
  if (prec == 0)
    hipLaunchKernelGGL(k_lu_front64<double>, g, b, 0, stream, fronts, flist, n_fronts, static_cast<double*>(arena),
                       piv, info);
  else if (prec == 1)
    hipLaunchKernelGGL(k_lu_front64<float>, g, b, 0, stream, fronts, flist, n_fronts, static_cast<float*>(arena),
                       piv, info);
  else
    return -1006;
  return ret(hipGetLastError());
}


int uhmk_lu_rows(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, void* arena,
                 int prec, const int* pperm, int width, hipStream_t stream) {
  if (n_items <= 0) return 0;
  if (nb > TILE || nb < 1) return -1000;
  if (prec == 0)
    return launch_rows<double>(fronts, items, n_items, c0, nb, static_cast<double*>(arena), pperm, width, stream);
  if (prec == 1)
    return launch_rows<float>(fronts, items, n_items, c0, nb, static_cast<float*>(arena), pperm, width, stream);
  return -1006;
}

}  // extern "C"
