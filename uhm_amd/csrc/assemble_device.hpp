// Extend-add of one 64 x 64 tile of a front in gather form (gfx950), shared
// by the assembly launch and the fused small-front Cholesky launch.
//
// Reference operation restated (semantics, not code): Helper_::_merge_A +
// Matrix_FLA_::merge, UHM/include/uhm/matrix/uhm/helper.hxx:191-222,
// UHM/src/matrix/uhm/fla/matrix.cxx:284-303; op_merge,
// UHM/src/operation/element.cxx:135-185.
#pragma once

#include <hip/hip_runtime.h>

#include "dense_device.hpp"
#include "uhmk.h"

namespace uhmd {

// Sources of one front: its own stage and the children's ABR blocks, in
// merge order (at most ASM_MAX_SRC).
constexpr int ASM_MAX_SRC = 16;
struct AsmSources {
  long long off[ASM_MAX_SRC];
  int ld[ASM_MAX_SRC], inv[ASM_MAX_SRC], kind[ASM_MAX_SRC];
  int ns;
};

// Source table of front f into S (threads 0 .. ns-1 write; the caller
// synchronises the workgroup before asm_tile reads it).
__device__ __forceinline__ void asm_sources(AsmSources& S, const uhmk_front* __restrict__ F, int f,
                                            const int* __restrict__ src_beg, const uhmk_source* __restrict__ srcs) {
  const int s0 = src_beg[f], ns = min(src_beg[f + 1] - s0, ASM_MAX_SRC);
  const int n = F[f].n;
  if (threadIdx.x == 0) S.ns = ns;
  if ((int)threadIdx.x < ns) {
    const uhmk_source sc = srcs[s0 + threadIdx.x];
    if (sc.src < 0) {
      S.kind[threadIdx.x] = 0;
      S.off[threadIdx.x] = F[f].soff;
      S.ld[threadIdx.x] = n;
      S.inv[threadIdx.x] = -1;
    } else {
      const uhmk_front ch = F[sc.src];
      S.kind[threadIdx.x] = 1 + ch.prec;
      S.off[threadIdx.x] = ch.off + ch.fs + (long long)ch.fs * ch.n;
      S.ld[threadIdx.x] = ch.n;
      S.inv[threadIdx.x] = sc.inv;
    }
  }
}

// Every entry of the 64 x 64 tile (i0, j0) of an n x n front at element off
// of its arena (T: FP64 or FP32) rebuilt as 0 + own stage + children's ABR,
// the sources summed in merge order in FP64 and rounded once to T -- the
// exact operation sequence of op_merge on a freshly zeroed front.  256
// threads: thread (tx, ty) takes row i0 + tx and columns j0 + ty + 4 u
// (u < 16); each source's 16 index and 16 value loads issue together.
template <typename T>
__device__ __forceinline__ void asm_tile(const AsmSources& S, int n, long long off, int i0, int j0,
                                         const int* __restrict__ inv, double* __restrict__ arena64,
                                         float* __restrict__ arena32, const double* __restrict__ stage) {
  constexpr int TL = 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int row = i0 + tx;
  if (row >= n) return;
  double acc[TL / 4];
#pragma unroll
  for (int u = 0; u < TL / 4; ++u) acc[u] = 0.0;
  for (int s = 0; s < S.ns; ++s) {
    const int k = S.kind[s];
    const int ri = k ? inv[S.inv[s] + row] : row;
    if (ri < 0) continue;
    const long long base = S.off[s] + ri;
    const int ld = S.ld[s];
    int cj[TL / 4];
#pragma unroll
    for (int u = 0; u < TL / 4; ++u) {
      const int col = j0 + ty + 4 * u;
      const bool ok = col < n;
      cj[u] = k ? (ok ? inv[S.inv[s] + col] : -1) : (ok ? col : -1);
    }
    if (k == 0) {
#pragma unroll
      for (int u = 0; u < TL / 4; ++u) acc[u] += ldz(stage, base + (long long)cj[u] * ld, cj[u] >= 0);
    } else if (k == 1) {
#pragma unroll
      for (int u = 0; u < TL / 4; ++u) acc[u] += ldz(arena64, base + (long long)cj[u] * ld, cj[u] >= 0);
    } else {
#pragma unroll
      for (int u = 0; u < TL / 4; ++u) acc[u] += ldz(arena32, base + (long long)cj[u] * ld, cj[u] >= 0);
    }
  }
  T* A = (sizeof(T) == 8 ? (T*)arena64 : (T*)arena32) + off;
#pragma unroll
  for (int u = 0; u < TL / 4; ++u) {
    const int col = j0 + ty + 4 * u;
    if (col < n) A[row + (long long)col * n] = (T)acc[u];
  }
}

}  // namespace uhmd
