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

#include "uhmk.h"

namespace {

constexpr int BLK = 64;

// ------------------------------------------------------------------------
// Forward solve of one right-hand-side column for every front in flist.
// x_f = b_f + sum_children xb_c (children gathered through inv, in merge
// order), then P (LU), then elimination with the unit (LU) or non-unit
// (Cholesky) lower factor over all n rows, which also forms
// xb -= ABL * xt.  Blocks of 64 columns: the diagonal triangle is solved by
// one wavefront with shuffles, the rows below by the whole workgroup.
// ------------------------------------------------------------------------
template <bool CHOL>
__global__ __launch_bounds__(256) void k_solve_fwd(const uhmk_front* __restrict__ F,
                                                   const int* __restrict__ flist,
                                                   const int* __restrict__ src_beg,
                                                   const uhmk_source* __restrict__ srcs,
                                                   const int* __restrict__ inv,
                                                   const double* __restrict__ arena,
                                                   const int* __restrict__ piv, double* __restrict__ x,
                                                   const double* __restrict__ b, int col) {
  extern __shared__ double sx[];
  const int f = flist[blockIdx.x];
  const uhmk_front fr = F[f];
  const int n = fr.n, fs = fr.fs;
  const int tid = threadIdx.x, lane = tid & 63;
  const long long xo = fr.xoff + (long long)col * n;
  const int s0 = src_beg[f], s1 = src_beg[f + 1];
  for (int i = tid; i < n; i += 256) {
    double acc = b[xo + i];
    for (int s = s0; s < s1; ++s) {
      const uhmk_source sc = srcs[s];
      if (sc.src < 0) continue;
      const int ri = inv[sc.inv + i];
      if (ri >= 0) {
        const uhmk_front ch = F[sc.src];
        acc += x[ch.xoff + (long long)col * ch.n + ch.fs + ri];
      }
    }
    sx[i] = acc;
  }
  __syncthreads();
  if (!CHOL && tid == 0) {
    for (int k = 0; k < fs; ++k) {
      const int p = piv[fr.poff + k];
      if (p) {
        const double t = sx[k];
        sx[k] = sx[k + p];
        sx[k + p] = t;
      }
    }
  }
  __syncthreads();
  const double* __restrict__ A = arena + fr.off;
  for (int k0 = 0; k0 < fs; k0 += BLK) {
    const int kb = min(BLK, fs - k0);
    if (tid < 64) {
      const int j = lane;
      double xj = (j < kb) ? sx[k0 + j] : 0.0;
      double Lr[BLK];
#pragma unroll
      for (int k = 0; k < BLK; ++k)
        Lr[k] = (j < kb && k < kb && k <= j) ? A[(k0 + j) + (long long)(k0 + k) * n] : 0.0;
#pragma unroll
      for (int k = 0; k < BLK; ++k) {
        if (k < kb) {
          if (CHOL && j == k) xj = xj / Lr[k];
          const double xk = __shfl(xj, k);
          if (j > k) xj -= Lr[k] * xk;
        }
      }
      if (j < kb) sx[k0 + j] = xj;
    }
    __syncthreads();
    for (int i = k0 + kb + tid; i < n; i += 256) {
      double acc = 0.0;
#pragma unroll 8
      for (int k = 0; k < kb; ++k) acc += A[i + (long long)(k0 + k) * n] * sx[k0 + k];
      sx[i] -= acc;
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += 256) x[xo + i] = sx[i];
}

// Backward solve (LU): xb <- parent's solution (branch), xt -= ATR * xb,
// xt <- U^-1 xt.
__global__ __launch_bounds__(256) void k_solve_bwd_lu(const uhmk_front* __restrict__ F,
                                                      const int* __restrict__ flist,
                                                      const int* __restrict__ emap,
                                                      const double* __restrict__ arena,
                                                      double* __restrict__ x, int col) {
  extern __shared__ double sx[];
  const int f = flist[blockIdx.x];
  const uhmk_front fr = F[f];
  const int n = fr.n, fs = fr.fs, ss = n - fs;
  const int tid = threadIdx.x, lane = tid & 63;
  const long long xo = fr.xoff + (long long)col * n;
  for (int i = tid; i < fs; i += 256) sx[i] = x[xo + i];
  if (fr.parent >= 0) {
    const uhmk_front P = F[fr.parent];
    for (int i = tid; i < ss; i += 256) sx[fs + i] = x[P.xoff + (long long)col * P.n + emap[fr.emap + i]];
  } else {
    for (int i = tid; i < ss; i += 256) sx[fs + i] = x[xo + fs + i];
  }
  __syncthreads();
  const double* __restrict__ A = arena + fr.off;
  for (int i = tid; i < fs; i += 256) {
    double acc = 0.0;
#pragma unroll 8
    for (int k = 0; k < ss; ++k) acc += A[i + (long long)(fs + k) * n] * sx[fs + k];
    sx[i] -= acc;
  }
  __syncthreads();
  const int nblk = (fs + BLK - 1) / BLK;
  for (int bk = nblk - 1; bk >= 0; --bk) {
    const int k0 = bk * BLK, kb = min(BLK, fs - k0);
    if (tid < 64) {
      const int j = lane;
      double xj = (j < kb) ? sx[k0 + j] : 0.0;
      double Ur[BLK];
#pragma unroll
      for (int k = 0; k < BLK; ++k)
        Ur[k] = (j < kb && k < kb && k >= j) ? A[(k0 + j) + (long long)(k0 + k) * n] : 0.0;
#pragma unroll
      for (int k = BLK - 1; k >= 0; --k) {
        if (k < kb) {
          if (j == k) xj = xj / Ur[k];
          const double xk = __shfl(xj, k);
          if (j < k) xj -= Ur[k] * xk;
        }
      }
      if (j < kb) sx[k0 + j] = xj;
    }
    __syncthreads();
    for (int i = tid; i < k0; i += 256) {
      double acc = 0.0;
#pragma unroll 8
      for (int k = 0; k < kb; ++k) acc += A[i + (long long)(k0 + k) * n] * sx[k0 + k];
      sx[i] -= acc;
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += 256) x[xo + i] = sx[i];
}

// Backward solve (Cholesky): xb <- parent's solution, xt -= ABL^T xb,
// xt <- L^-T xt.
__global__ __launch_bounds__(256) void k_solve_bwd_chol(const uhmk_front* __restrict__ F,
                                                        const int* __restrict__ flist,
                                                        const int* __restrict__ emap,
                                                        const double* __restrict__ arena,
                                                        double* __restrict__ x, int col) {
  extern __shared__ double sx[];
  const int f = flist[blockIdx.x];
  const uhmk_front fr = F[f];
  const int n = fr.n, fs = fr.fs, ss = n - fs;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const long long xo = fr.xoff + (long long)col * n;
  for (int i = tid; i < fs; i += 256) sx[i] = x[xo + i];
  if (fr.parent >= 0) {
    const uhmk_front P = F[fr.parent];
    for (int i = tid; i < ss; i += 256) sx[fs + i] = x[P.xoff + (long long)col * P.n + emap[fr.emap + i]];
  } else {
    for (int i = tid; i < ss; i += 256) sx[fs + i] = x[xo + fs + i];
  }
  __syncthreads();
  const double* __restrict__ A = arena + fr.off;
  for (int i = wid; i < fs; i += 4) {
    double acc = 0.0;
#pragma unroll 8
    for (int k = lane; k < ss; k += 64) acc += A[(fs + k) + (long long)i * n] * sx[fs + k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) sx[i] -= acc;
  }
  __syncthreads();
  const int nblk = (fs + BLK - 1) / BLK;
  for (int bk = nblk - 1; bk >= 0; --bk) {
    const int k0 = bk * BLK, kb = min(BLK, fs - k0), k1 = k0 + kb;
    for (int c = wid; c < kb; c += 4) {
      double acc = 0.0;
#pragma unroll 8
      for (int i = k1 + lane; i < fs; i += 64) acc += A[i + (long long)(k0 + c) * n] * sx[i];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
      if (lane == 0) sx[k0 + c] -= acc;
    }
    __syncthreads();
    if (tid < 64) {
      const int j = lane;
      double xj = (j < kb) ? sx[k0 + j] : 0.0;
      double Lc[BLK];
#pragma unroll
      for (int i = 0; i < BLK; ++i)
        Lc[i] = (j < kb && i < kb && i >= j) ? A[(k0 + i) + (long long)(k0 + j) * n] : 0.0;
#pragma unroll
      for (int i = BLK - 1; i >= 0; --i) {
        if (i < kb) {
          if (j == i) xj = xj / Lc[i];
          const double xi = __shfl(xj, i);
          if (j < i) xj -= Lc[i] * xi;
        }
      }
      if (j < kb) sx[k0 + j] = xj;
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += 256) x[xo + i] = sx[i];
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
    for (int idx = tid; idx < m * m; idx += 256) {
      const int i = idx % m, j = idx / m;
      double acc = 0.0;
      for (int k = 0; k < m; ++k) acc += sM[k + i * ldm] * sM[k + j * ldm];
      if (i == j) acc += alpha;
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

// ------------------------------------------------------------------------
// Assembled-matrix residual r = b - A x (CSR, FP64) and the norms of the
// normwise backward error ||r||inf / (||A||inf ||x||inf + ||b||inf).
// out[0..3] += {max|r_i|, max row sum |A|, max|x_i|, max|b_i|} by atomics on
// non-negative doubles ordered as unsigned 64-bit integers.
// ------------------------------------------------------------------------
__device__ inline void atomic_max_nonneg(double* p, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

__global__ __launch_bounds__(256) void k_residual(int nrow, const long long* __restrict__ rowptr,
                                                  const int* __restrict__ colidx,
                                                  const double* __restrict__ val,
                                                  const double* __restrict__ xg,
                                                  const double* __restrict__ bg, double* __restrict__ r,
                                                  double* __restrict__ out) {
  // one wavefront per row, grid-stride over rows; maxima kept per wave and
  // published with one atomic each at the end
  const int lane = threadIdx.x & 63;
  const int nwaves = gridDim.x * 4;
  double mr = 0.0, ma = 0.0, mx = 0.0, mb = 0.0;
  for (int row = blockIdx.x * 4 + (threadIdx.x >> 6); row < nrow; row += nwaves) {
    double ax = 0.0, an = 0.0;
    for (long long p = rowptr[row] + lane; p < rowptr[row + 1]; p += 64) {
      const double a = val[p];
      ax += a * xg[colidx[p]];
      an += fabs(a);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ax += __shfl_xor(ax, o);
      an += __shfl_xor(an, o);
    }
    const double ri = bg[row] - ax;
    if (lane == 0) r[row] = ri;
    mr = fmax(mr, fabs(ri));
    ma = fmax(ma, an);
    mx = fmax(mx, fabs(xg[row]));
    mb = fmax(mb, fabs(bg[row]));
  }
  if (lane == 0) {
    atomic_max_nonneg(out + 0, mr);
    atomic_max_nonneg(out + 1, ma);
    atomic_max_nonneg(out + 2, mx);
    atomic_max_nonneg(out + 3, mb);
  }
}

// Global <-> front vectors: g[gidx[k]] = xf[k] (dir 0) or xf[k] = g[gidx[k]]
// (dir 1) over the factor rows of every front; item list = fronts.
__global__ __launch_bounds__(256) void k_front_global(const uhmk_front* __restrict__ F, int nfront,
                                                      const long long* __restrict__ goff,
                                                      double* __restrict__ xf, double* __restrict__ g,
                                                      int col, long long nglob, int dir) {
  const int f = blockIdx.x;
  if (f >= nfront) return;
  const uhmk_front fr = F[f];
  const long long xo = fr.xoff + (long long)col * fr.n;
  const long long go = goff[f] + (long long)col * nglob;
  for (int i = threadIdx.x; i < fr.fs; i += 256) {
    if (dir == 0) g[go + i] = xf[xo + i];
    else xf[xo + i] = g[go + i];
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

inline int ret(hipError_t e) { return e == hipSuccess ? 0 : -(int)e; }

}  // namespace

extern "C" {

int uhmk_solve_fwd(const uhmk_front* fronts, const int* flist, int n_fronts, int max_n, int chol,
                   const int* src_beg, const uhmk_source* srcs, const int* inv, const double* arena,
                   const int* piv, double* x, const double* b, int col, hipStream_t stream) {
  if (n_fronts <= 0) return 0;
  const size_t sh = (size_t)max_n * sizeof(double);
  if (sh > 160 * 1024) return -1002;
  if (chol)
    hipLaunchKernelGGL(k_solve_fwd<true>, dim3(n_fronts), dim3(256), sh, stream, fronts, flist, src_beg, srcs,
                       inv, arena, piv, x, b, col);
  else
    hipLaunchKernelGGL(k_solve_fwd<false>, dim3(n_fronts), dim3(256), sh, stream, fronts, flist, src_beg, srcs,
                       inv, arena, piv, x, b, col);
  return ret(hipGetLastError());
}

int uhmk_solve_bwd(const uhmk_front* fronts, const int* flist, int n_fronts, int max_n, int chol,
                   const int* emap, const double* arena, double* x, int col, hipStream_t stream) {
  if (n_fronts <= 0) return 0;
  const size_t sh = (size_t)max_n * sizeof(double);
  if (sh > 160 * 1024) return -1002;
  if (chol)
    hipLaunchKernelGGL(k_solve_bwd_chol, dim3(n_fronts), dim3(256), sh, stream, fronts, flist, emap, arena, x, col);
  else
    hipLaunchKernelGGL(k_solve_bwd_lu, dim3(n_fronts), dim3(256), sh, stream, fronts, flist, emap, arena, x, col);
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

int uhmk_residual(int nrow, const long long* rowptr, const int* colidx, const double* val, const double* xg,
                  const double* bg, double* r, double* out4, hipStream_t stream) {
  if (nrow <= 0) return 0;
  const int blocks = (nrow + 3) / 4 < 2048 ? (nrow + 3) / 4 : 2048;
  hipLaunchKernelGGL(k_residual, dim3(blocks), dim3(256), 0, stream, nrow, rowptr, colidx, val, xg, bg, r, out4);
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

int uhmk_front_global(const uhmk_front* fronts, int n_fronts, const long long* goff, double* xf, double* g,
                      int col, long long nglob, int dir, hipStream_t stream) {
  if (n_fronts <= 0) return 0;
  hipLaunchKernelGGL(k_front_global, dim3(n_fronts), dim3(256), 0, stream, fronts, n_fronts, goff, xf, g, col,
                     nglob, dir);
  return ret(hipGetLastError());
}

}  // extern "C"
