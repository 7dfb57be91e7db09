// Device helpers shared by the dense front kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

namespace uhmd {

// Triangular solve on 64-long vectors, 16 vectors per wavefront: lane
// l = v + 16 q holds entries [16 q, 16 q + 16) of vector v in x[0..16).
// For k = 0 .. kb-1 in order:
//     x_k <- x_k * rdiag[k]                 (skipped when UNIT)
//     x_c <- x_c - x_k * T(c, k)   for k < c < kb
// with T(c, k) = s[c * ld + k] (TRANS = false) or s[k * ld + c] (TRANS = true)
// read from LDS.  This is the right-looking (column) form of:
//   * X <- X L^-T          rows of X, T = L   (Cholesky panel rows)
//   * X <- X U^-1          rows of X, T = U^T (LU rows of ABL: TRANS)
//   * x <- L^-1 x, unit    columns,   T = L   (LU U strip)
// Every vector is independent; the owner of x_k broadcasts it to the other
// three lanes of its vector with a shuffle, so no barrier is needed.
template <bool UNIT, bool TRANS>
__device__ inline void tri_solve64(double (&x)[16], const double* __restrict__ s, int ld,
                                   const double* __restrict__ rdiag, int kb, int lane) {
  const int q = lane >> 4, v = lane & 15;
#pragma unroll
  for (int k = 0; k < 64; ++k) {
    if (k < kb) {
      const int kq = k >> 4, kr = k & 15;
      if (!UNIT && q == kq) x[kr] *= rdiag[k];
      const double xk = __shfl(x[kr], v + 16 * kq);
#pragma unroll
      for (int cr = 0; cr < 16; ++cr) {
        const int c = 16 * q + cr;
        if (c > k && c < kb) x[cr] -= xk * (TRANS ? s[k * ld + c] : s[c * ld + k]);
      }
    }
  }
}

}  // namespace uhmd
