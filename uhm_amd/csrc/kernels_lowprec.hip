// Low-precision Schur complement updates for gfx950: the operands of
// ABR -= ABL ATR (LU) / ABR -= ABL ABL^T (Cholesky) packed once per front
// into BF16 or FP16 (nearest even; FP16 with a power-of-two scale per
// operand and front, so that the largest entry lies in [2^14, 2^15)), then
// multiplied on v_mfma_f32_16x16x32_{bf16,f16} with FP32 accumulation, the
// front updated in its storage type.
//
// Reference operations restated (semantics, not code): the final Schur
// complement of lu_piv_hier / chol_hier,
//   UHM/src/matrix/uhm/fla/lu_piv/decompose.cxx:148-149 (gemm)
//   UHM/src/matrix/uhm/fla/chol/decompose.cxx:95-96    (syrk)
//   LINAL/src/dense/gemm.cxx:48-137, LINAL/src/dense/syrk.cxx:43-123
//
// Packed layout (per front f, element offset off[f] in the packed buffer):
// operand A = ABL as rows [ss_pad][ldk] (row i = ABL row i, k contiguous),
// for LU operand B = ATR^T at off[f] + ss_pad * ldk (row j = ATR column
// j); ldk = fs rounded up to 64, ss_pad = ss rounded up to 128, padding
// zero, so the product kernel reads whole 128 x 64 tiles without bounds
// checks.  Cholesky uses A for both operands.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dense_device.hpp"
#include "uhmk.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

constexpr int PK = 64;  // pack tile edge

// Largest |entry| of every pack tile into amax[2 f + op] (the bits of a
// non-negative float order as unsigned ints).  items {f, row tile, k tile, op}.
template <typename TS>
__global__ __launch_bounds__(256) void k_lp_absmax(const uhmk_front* __restrict__ F,
                                                   const uhmk_item* __restrict__ items, const TS* __restrict__ arena,
                                                   unsigned* __restrict__ amax) {
  __shared__ float red[4];
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs, ss = n - fs;
  const TS* __restrict__ M = arena + fr.off;
  const int tid = threadIdx.x, r0 = PK * it.a, k0 = PK * it.b, op = it.c;
  float m = 0.0f;
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int e = tid + 256 * u;
    // op 0 (ABL, column major): consecutive threads run down a column; op 1
    // (ATR^T): along k of one ATR column
    const int row = r0 + (op ? e >> 6 : e & 63), k = k0 + (op ? e & 63 : e >> 6);
    if (row < ss && k < fs) {
      const double v = op ? (double)M[k + (long long)(fs + row) * n] : (double)M[(fs + row) + (long long)k * n];
      m = fmaxf(m, (float)fabs(v));
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if ((tid & 63) == 0) red[tid >> 6] = m;
  __syncthreads();
  if (tid == 0) {
    m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    if (m > 0.0f) atomicMax(amax + 2 * it.f + op, __float_as_uint(m));
  }
}

// Power-of-two scale of an FP16 operand: amax * s in [2^14, 2^15) (1 for
// BF16 or an all-zero operand).
__device__ __forceinline__ float lp_scale(int fp16, unsigned amax_bits) {
  if (!fp16 || amax_bits == 0) return 1.0f;
  int e;
  (void)frexpf(__uint_as_float(amax_bits), &e);  // amax < 2^e
  return ldexpf(1.0f, 15 - e);
}

// One 64 x 64 tile of the packed operand (zero outside the front), through
// LDS so that both the reads of the front and the writes of packed rows
// are coalesced.  items {f, row tile, k tile, op}.
template <typename TS, int FP16>
__global__ __launch_bounds__(256) void k_lp_pack(const uhmk_front* __restrict__ F,
                                                 const uhmk_item* __restrict__ items, const TS* __restrict__ arena,
                                                 const long long* __restrict__ off, const unsigned* __restrict__ amax,
                                                 unsigned short* __restrict__ lp) {
  __shared__ double sT[PK][PK + 1];  // [row][k], the scaled entries (d * s is exact)
  const uhmk_item it = items[blockIdx.x];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs, ss = n - fs;
  const TS* __restrict__ M = arena + fr.off;
  const int tid = threadIdx.x, r0 = PK * it.a, k0 = PK * it.b, op = it.c;
  const int ldk = (fs + 63) & ~63, ss_pad = (ss + 127) & ~127;
  const float s = lp_scale(FP16, amax[2 * it.f + op]);
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int e = tid + 256 * u;
    const int rr = op ? e >> 6 : e & 63, kk = op ? e & 63 : e >> 6;
    const int row = r0 + rr, k = k0 + kk;
    double v = 0.0;
    if (row < ss && k < fs) {
      const double d = op ? (double)M[k + (long long)(fs + row) * n] : (double)M[(fs + row) + (long long)k * n];
      v = d * (double)s;
    }
    sT[rr][kk] = v;
  }
  __syncthreads();
  unsigned short* __restrict__ dst = lp + off[it.f] + (op ? (long long)ss_pad * ldk : 0);
  // 64 rows x 8 vectors of eight entries: two per thread
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int v8 = tid + 256 * u, rr = v8 >> 3, q = v8 & 7;
    u16x8 out;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      // one rounding to the packed type, from the storage value itself (an
      // FP64 entry is not rounded to FP32 first)
      const double x = sT[rr][8 * q + e];
      if (FP16) {
        const _Float16 h = (_Float16)x;
        out[e] = __builtin_bit_cast(unsigned short, h);
      } else {
        const __bf16 b = (__bf16)x;
        out[e] = __builtin_bit_cast(unsigned short, b);
      }
    }
    *reinterpret_cast<u16x8*>(dst + (long long)(r0 + rr) * ldk + k0 + 8 * q) = out;
  }
}

template <int FP16>
struct LpMma;
template <>
struct LpMma<0> {
  typedef bf16x8 v_t;
  __device__ __forceinline__ static f4 run(v_t a, v_t b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
template <>
struct LpMma<1> {
  typedef f16x8 v_t;
  __device__ __forceinline__ static f4 run(v_t a, v_t b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

// C -= A B^T on T x T tiles of ABR (T = 64 or 128), A and B packed rows
// (B = A for Cholesky), K = ldk in 32-deep stages double-buffered in LDS
// (rows padded to 40 entries), one barrier per stage; four wavefronts each
// own a (T/2)^2 quarter of 16 x 16 blocks.  With T = 64 a workgroup holds
// 20 KB of LDS and 52 VGPRs per lane, so eight workgroups share a CU: on a
// 4225 x 4225 x 2080 product this ran 2.6x faster than the same tile in
// 64-deep stages (four per CU) and 2x faster than 128-wide tiles, whose
// ~600 tiles leave most CUs with two waves per SIMD; neither a swizzled
// unpadded LDS layout nor loads issued two stages ahead changed it.  The product is
// formed transposed (D = B A^T) so that a 16-lane group stores 16
// consecutive rows of C.  items {f, i0, j0, 0} (front indices >= fs).
template <int T, typename TS, int FP16, bool LU>
__global__ __launch_bounds__(256, 2) void k_schur_lp(const uhmk_front* __restrict__ F,
                                                  const uhmk_item* __restrict__ items, TS* __restrict__ arena,
                                                  const long long* __restrict__ off,
                                                  const unsigned* __restrict__ amax,
                                                  const unsigned short* __restrict__ lp) {
  typedef typename LpMma<FP16>::v_t v_t;
  constexpr int KS = 32, LDK = KS + 8, NB = T / 32;
  constexpr int TPR = KS / 8;       // threads per packed row segment (8 entries each)
  constexpr int RPT = T * TPR / 256;  // rows per thread and operand per stage
  __shared__ __align__(16) unsigned short sA[2][T * LDK];
  __shared__ __align__(16) unsigned short sB[2][T * LDK];
  const uhmk_item it = items[uhmd::xcd_item(blockIdx.x, gridDim.x)];
  const uhmk_front fr = F[it.f];
  const int n = fr.n, fs = fr.fs, ss = n - fs;
  const int ldk = (fs + 63) & ~63, ss_pad = (ss + 127) & ~127;
  const int i0 = it.a, j0 = it.b;
  const unsigned short* __restrict__ pa = lp + off[it.f] + (long long)(i0 - fs) * ldk;
  const unsigned short* __restrict__ pb =
      lp + off[it.f] + (LU ? (long long)ss_pad * ldk : 0) + (long long)(j0 - fs) * ldk;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int iw = (w & 1) * (T / 2), jw = (w >> 1) * (T / 2);
  const int lr = tid / TPR, lq = (tid % TPR) * 8;
  constexpr int RS = 256 / TPR;  // row step between a thread's rows

  f4 acc[NB][NB];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = f4{0, 0, 0, 0};

  auto load = [&](uint4 (&ra)[RPT], uint4 (&rb)[RPT], int k0) {
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
      ra[u] = *reinterpret_cast<const uint4*>(pa + (long long)(lr + RS * u) * ldk + k0 + lq);
      rb[u] = *reinterpret_cast<const uint4*>(pb + (long long)(lr + RS * u) * ldk + k0 + lq);
    }
  };
  auto store = [&](const uint4 (&ra)[RPT], const uint4 (&rb)[RPT], int buf) {
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
      *reinterpret_cast<uint4*>(&sA[buf][(lr + RS * u) * LDK + lq]) = ra[u];
      *reinterpret_cast<uint4*>(&sB[buf][(lr + RS * u) * LDK + lq]) = rb[u];
    }
  };
  const int cl = lane & 15, kq = 8 * (lane >> 4);
  auto compute = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < KS; kk += 32) {
      v_t ao[NB], bo[NB];
#pragma unroll
      for (int a = 0; a < NB; ++a) ao[a] = *reinterpret_cast<const v_t*>(&sA[buf][(iw + a * 16 + cl) * LDK + kk + kq]);
#pragma unroll
      for (int b = 0; b < NB; ++b) bo[b] = *reinterpret_cast<const v_t*>(&sB[buf][(jw + b * 16 + cl) * LDK + kk + kq]);
#pragma unroll
      for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = LpMma<FP16>::run(bo[b], ao[a], acc[a][b]);
    }
  };
  // stage st + 1 in flight in registers while stage st is multiplied
  const int nst = ldk / KS;
  uint4 ra[RPT], rb[RPT];
  if (nst > 0) {
    load(ra, rb, 0);
    store(ra, rb, 0);
  }
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    if (st + 1 < nst) load(ra, rb, (st + 1) * KS);
    compute(st & 1);
    if (st + 1 < nst) store(ra, rb, (st & 1) ^ 1);
    __syncthreads();
  }
  // undo the operand scales (powers of two: exact)
  const double inv = 1.0 / ((double)lp_scale(FP16, amax[2 * it.f]) * (double)lp_scale(FP16, amax[2 * it.f + (LU ? 1 : 0)]));
  TS* __restrict__ M = arena + fr.off;
  const int i1 = min(i0 + T, n), j1 = min(j0 + T, n);
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int i = i0 + iw + a * 16 + cl;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = j0 + jw + b * 16 + 4 * (lane >> 4) + r;
        if (i < i1 && j < j1) {
          TS* c = M + i + (long long)j * n;
          *c = (TS)((double)*c - (double)acc[a][b][r] * inv);
        }
      }
    }
}

template <typename TS, int FP16>
int launch_schur_lp(const uhmk_front* fronts, const uhmk_item* items, int n_items, int lu, int tile, TS* arena,
                    const long long* off, const unsigned* amax, const unsigned short* lp, hipStream_t stream) {
  dim3 g(n_items), b(256);
  if (tile == 128) {
    if (lu) hipLaunchKernelGGL((k_schur_lp<128, TS, FP16, true>), g, b, 0, stream, fronts, items, arena, off, amax, lp);
    else hipLaunchKernelGGL((k_schur_lp<128, TS, FP16, false>), g, b, 0, stream, fronts, items, arena, off, amax, lp);
  } else if (tile == 64) {
    if (lu) hipLaunchKernelGGL((k_schur_lp<64, TS, FP16, true>), g, b, 0, stream, fronts, items, arena, off, amax, lp);
    else hipLaunchKernelGGL((k_schur_lp<64, TS, FP16, false>), g, b, 0, stream, fronts, items, arena, off, amax, lp);
  } else {
    return -1004;
  }
  return 0;
}

inline int ret(hipError_t e) { return e == hipSuccess ? 0 : -(int)e; }

}  // namespace

extern "C" {

long long uhmk_schur_lp_size(int n, int fs, int lu) {
  const long long ss = n - fs;
  if (fs <= 0 || ss <= 0) return 0;
  const long long ldk = (fs + 63) & ~63LL, ss_pad = (ss + 127) & ~127LL;
  return (lu ? 2 : 1) * ss_pad * ldk;
}

int uhmk_schur_lp_pack(const uhmk_front* fronts, const uhmk_item* items, int n_items, int fp16, const void* arena,
                       int prec, const long long* off, unsigned* amax, unsigned short* lp, hipStream_t stream) {
  if (n_items <= 0) return 0;
  dim3 g(n_items), b(256);
  if (prec == 0) {
    const double* A = static_cast<const double*>(arena);
    if (fp16) {
      hipLaunchKernelGGL(k_lp_absmax<double>, g, b, 0, stream, fronts, items, A, amax);
      hipLaunchKernelGGL((k_lp_pack<double, 1>), g, b, 0, stream, fronts, items, A, off, amax, lp);
    } else {
      hipLaunchKernelGGL((k_lp_pack<double, 0>), g, b, 0, stream, fronts, items, A, off, amax, lp);
    }
  } else if (prec == 1) {
    const float* A = static_cast<const float*>(arena);
    if (fp16) {
      hipLaunchKernelGGL(k_lp_absmax<float>, g, b, 0, stream, fronts, items, A, amax);
      hipLaunchKernelGGL((k_lp_pack<float, 1>), g, b, 0, stream, fronts, items, A, off, amax, lp);
    } else {
      hipLaunchKernelGGL((k_lp_pack<float, 0>), g, b, 0, stream, fronts, items, A, off, amax, lp);
    }
  } else {
    return -1006;
  }
  return ret(hipGetLastError());
}

int uhmk_schur_lp(const uhmk_front* fronts, const uhmk_item* items, int n_items, int lu, int tile, int fp16,
                  void* arena, int prec, const long long* off, const unsigned* amax, const unsigned short* lp,
                  hipStream_t stream) {
  if (n_items <= 0) return 0;
  int rc;
  if (prec == 0)
    rc = fp16 ? launch_schur_lp<double, 1>(fronts, items, n_items, lu, tile, static_cast<double*>(arena), off, amax,
                                           lp, stream)
              : launch_schur_lp<double, 0>(fronts, items, n_items, lu, tile, static_cast<double*>(arena), off, amax,
                                           lp, stream);
  else if (prec == 1)
    rc = fp16 ? launch_schur_lp<float, 1>(fronts, items, n_items, lu, tile, static_cast<float*>(arena), off, amax, lp,
                                          stream)
              : launch_schur_lp<float, 0>(fronts, items, n_items, lu, tile, static_cast<float*>(arena), off, amax, lp,
                                          stream);
  else
    return -1006;
  if (rc) return rc;
  return ret(hipGetLastError());
}

}  // extern "C"
