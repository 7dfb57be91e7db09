/*
 * uhmk.h -- kernel-level C ABI of the MI355X multifrontal path.
 *
 * Every entry point launches hand-written HIP kernels (gfx950) on the
 * stream it is given and returns 0 on success or a negative HIP error code.
 * All pointers are device pointers; all calls are stream ordered and made
 * from one host thread.  These functions replace the per-front numeric seam
 * of the reference, where each front operation went through libflame
 * (LINAL/src/internal/gemm.cxx:78-106, LINAL/src/internal/trsm.cxx:70-97)
 * one b x b tile at a time; here one launch covers every front of a tree
 * level (a "batch"), bucketed by work item.
 *
 * Front layout (shared by every kernel): front f is one column-major
 * n x n matrix, n = fs + ss, leading dimension n, stored at arena + off in
 * the FP64 arena (prec 0) or the FP32 arena (prec 1); the fronts of one
 * tree level share a storage type.  Kernels that take `void* arena, int
 * prec` read and write that level's arena; arithmetic outside the MFMA
 * products is FP64.
 * Its quadrants are the reference's ATL (fs x fs), ABL (ss x fs),
 * ATR (fs x ss) and ABR (ss x ss) (UHM/src/matrix/uhm/fla/matrix.cxx:75-91).
 * Pivots are int32, libflame's relative format: p[k] = pivot row - k.
 */
#ifndef UHMK_H_
#define UHMK_H_

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct uhmk_front {
  long long off;   /* element offset of the front in its arena            */
  long long xoff;  /* element offset of the n x n_rhs block in x / b      */
  long long soff;  /* element offset of the assembly stage, -1 if none    */
  int n;           /* fs + ss                                             */
  int fs;          /* number of fully summed (factor) dofs                */
  int poff;        /* offset of the fs pivots in the pivot arena          */
  int parent;      /* parent front index, -1 for a root                   */
  int emap;        /* offset of the ss-long map child schur -> parent idx */
  int id;          /* element id, reported on a zero pivot / non-SPD      */
  int prec;        /* storage of the front: 0 FP64 arena, 1 FP32 arena   */
  int rsv;         /* reserved, 0                                         */
} uhmk_front;

/* Assembly source of a front: its own stage (src = -1) or the ABR block of
 * child front src, gathered through inv[0..n) (parent index -> child schur
 * index, -1 where the child does not contribute). */
typedef struct uhmk_source {
  int src;
  int inv;
} uhmk_source;

/* One copy-in block: an m x ncol element matrix (ncol = m, side 0) or
 * right-hand side block (side 1) of front f, column major at element koff
 * of the packed values, its rows' (and for side 0 columns') front positions
 * at element moff of the packed maps. */
typedef struct uhmk_cblock {
  long long koff;
  int moff, m, ncol, f, side;
} uhmk_cblock;

/* Work item: four ints whose meaning is fixed per kernel. */
typedef struct uhmk_item {
  int f, a, b, c;
} uhmk_item;

/* ---- factorization (one launch per tree level and panel step) ---------- */

/* Extend-add in gather form: every front in items is rebuilt as
 * 0 + own stage + children's ABR (merge order).  Replaces op_merge /
 * Helper_::merge_A / Matrix_FLA_::merge (UHM/src/operation/element.cxx:135-185,
 * UHM/include/uhm/matrix/uhm/helper.hxx:191-222,
 * UHM/src/matrix/uhm/fla/matrix.cxx:284-303).  items: {f, row tile, col tile}. */
int uhmk_assemble(const uhmk_front* fronts, const uhmk_item* items, int n_items, const int* src_beg,
                  const uhmk_source* srcs, const int* inv, double* arena64, float* arena32, int prec,
                  const double* stage, hipStream_t stream);

/* Element front-end of copy_in (interf.cxx:108-204 -> Helper_::merge_A /
 * merge_rhs_b into zeroed leaf fronts, UHM/include/uhm/matrix/uhm/
 * helper.hxx:191-331): one workgroup per item {f, first block, blocks,
 * side | zero_first << 1} adds its blocks, in order, into the stage of
 * leaf front f (side 0: stage[map[i] + map[j] n] += K[i + j m]) or its
 * right-hand side (side 1: b[xoff + map[i] + r n] += K[i + r m]), after
 * zeroing the n x n stage / n x nrhs block when zero_first is set; within
 * a block the positions are distinct, so the sums at every position are
 * formed in block order, as the host scatter forms them. */
int uhmk_copy_in(const uhmk_front* fronts, const uhmk_item* items, int n_items, const uhmk_cblock* blocks,
                 const double* vals, const int* maps, double* stage, double* b, int nrhs, hipStream_t stream);

/* count entries of a front from one arena into another, converted between
 * storage types (prec 0 FP64, 1 FP32): the factored front of an element
 * that a re-factorization reuses (Mesh_::check_reuse,
 * UHM/src/mesh/mesh.cxx:159-225), moved into the new pattern's arena. */
int uhmk_copy_front(const void* src, int src_prec, void* dst, int dst_prec, long long count, hipStream_t stream);

/* Panel LU with partial pivoting restricted to ATL rows [c0, fs) for the
 * fronts in flist; writes relative pivots.  Replaces FLA_LU_piv on the
 * panel inside linal::dense::lu_piv (LINAL/src/dense/lu_piv.cxx:79).
 * info[f] <- first zero-pivot row + 1.  pperm (257 ints per front, indexed
 * by front) receives the panel's net row permutation for uhmk_lu_rows:
 * count, then up to 128 rows, then up to 128 source list indices. */
int uhmk_lu_panel(const uhmk_front* fronts, const int* flist, int n_fronts, int max_rows, int c0, int nb,
                  void* arena, int prec, int* piv, int* info, int* pperm, hipStream_t stream);

/* The same panel (same arguments, results and pperm layout; the same bits
 * on FP64 fronts), left-looking over register blocks with the rows left in
 * place until the end: one barrier per column, the kb pivot rows held in
 * LDS.  Up to 8192 pivot-candidate rows (uhmk_lu_panel: 6144); -1001
 * above.  Replaces the same FLA_LU_piv call (LINAL/src/dense/lu_piv.cxx:79). */
int uhmk_lu_panel_ll(const uhmk_front* fronts, const int* flist, int n_fronts, int max_rows, int c0, int nb,
                     void* arena, int prec, int* piv, int* info, int* pperm, hipStream_t stream);

/* Row swaps outside the panel (apply_pivots, LINAL/src/dense/pivot.cxx:46-91)
 * from the panel permutation in pperm, U12 = L11^-1 A12 (trsm left lower
 * unit) and L21 = A21 U11^-1 on the ABL rows (trsm right upper,
 * UHM/src/matrix/uhm/fla/lu_piv/decompose.cxx:136-146).
 * items: {f, kind (0 left cols, 1 ABL rows, 2 right cols), start, 0}, each
 * a chunk of `width` columns (rows for kind 1): 16, 32 or 64. */
int uhmk_lu_rows(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, void* arena,
                 int prec, const int* pperm, int width, hipStream_t stream);

/* Whole-front LU of fronts with n <= 64, one wavefront per front, rows in
 * registers: pivots searched in ATL (relative format as uhmk_lu_panel),
 * L\U in ATL, U12 in ATR, L21 in ABL and the Schur complement in ABR in
 * one launch, replacing the panel / row / GEMM launches of lu_piv_hier
 * (UHM/src/matrix/uhm/fla/lu_piv/decompose.cxx:124-153) for such fronts.
 * info[f] = first zero-pivot row + 1. */
int uhmk_lu_front64(const uhmk_front* fronts, const int* flist, int n_fronts, void* arena, int prec, int* piv,
                    int* info, hipStream_t stream);

/* Cholesky of the diagonal panel block (FLA_Chol in linal::dense::chol,
 * LINAL/src/dense/chol.cxx:67) and its inverse W = L11^-1, whose strict
 * lower part is stored transposed in the strict upper triangle of the block
 * (never read as matrix data on the Cholesky path).  info[f] <- first
 * non-positive pivot + 1. */
int uhmk_chol_diag(const uhmk_front* fronts, const int* flist, int n_fronts, int c0, int nb, void* arena, int prec,
                   int* info, hipStream_t stream);

/* Panel rows below the diagonal block: X <- X L11^-T = X W^T with the
 * inverse from uhmk_chol_diag, as an FP64 MFMA product (trsm right lower
 * transpose, LINAL/src/dense/chol.cxx:80-84).  items: {f, 0, first row, 0}. */
int uhmk_chol_trsm(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, void* arena,
                   int prec, hipStream_t stream);

/* Cholesky panel update as uhmk_gemm_update (schur = 0, NT, tile 64,
 * nb = 64) in which the workgroup of each item flagged c = 1 -- the tile
 * (c1, c1), c1 = min(c0 + nb, fs), i.e. the next panel's diagonal block,
 * listed before the other items -- then factors that block as
 * uhmk_chol_diag(c0 = c1) does, from its updated values in LDS: the next
 * panel's diagonal launch is folded into this one (same results, bit for
 * bit).  mma 0 (FP64 fronts) or 1 (FP32 products); info as uhmk_chol_diag. */
int uhmk_chol_gemm_diag(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, void* arena,
                        int prec, int mma, int* info, hipStream_t stream);

/* Cholesky panel group in one persistent launch: the columns [c0, ge),
 * ge = min(c0 + ncol, fs), of every front listed, factored left-looking in
 * 64-column panels -- per panel the diagonal block and its inverse (as
 * uhmk_chol_diag), the panel rows below it (as uhmk_chol_trsm) and the
 * updates inside the group (as uhmk_gemm_update with K = 64 panels) -- with
 * the dependencies between panels resolved inside the launch (the steps of
 * linal::dense::chol, LINAL/src/dense/chol.cxx:46-92, for the columns of
 * one group; the trailing matrix right of the group is left to
 * uhmk_gemm_update).  The rows [c0, n) of a front are cut into 64-row
 * tiles (cut at fs too), numbered from 0; items {f, first tile r, number
 * of tiles, flag slot s}: each tile r < P (the panels of the group) an
 * item of its own, listed before the front's other items, which cover the
 * remaining tiles in increasing r.  flags: ints, zero before the
 * launch; flags[0] counts waits that gave up (nonzero: the results are
 * invalid), flags[s + r] the progress of tile r (at most 16 per front).
 * ncol a multiple of 64, at most 1024.  mma 0 (FP64 fronts) or 1 (FP32
 * products); info as uhmk_chol_diag. */
int uhmk_chol_group(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int ncol, void* arena,
                    int prec, int mma, int* flags, int* info, hipStream_t stream);

/* C -= A op(B) on tile x tile blocks (tile 64 or 128) on MFMA: panel
 * update (schur = 0, K = [c0, c0+nb)) or final Schur complement
 * ABR -= ABL ATR / ABL ABL^T (schur = 1, K = [0, fs)).  mma 0: FP64
 * products (v_mfma_f64_16x16x4f64, FP64 fronts only); mma 1: FP32
 * products and accumulation (v_mfma_f32_16x16x4f32) on FP32 or FP64
 * fronts; mma 2: A and B rounded to BF16, FP32 products and accumulation
 * (v_mfma_f32_16x16x32_bf16) on FP32 or FP64 fronts.  Replaces linal::dense::gemm / syrk (LINAL/src/dense/gemm.cxx:48-137,
 * syrk.cxx:43-123).  items: {f, i0, j0, 0}, tiles cut at fs. */
int uhmk_gemm_update(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, int schur,
                     int transb, int tile, void* arena, int prec, int mma, hipStream_t stream);

/* Whole Cholesky factorization of small fronts, one workgroup per front of
 * flist (every front needs fs <= 64; one with fs > 64 is left untouched and
 * reports info[f] = -1): the front is assembled as uhmk_assemble does, its
 * diagonal block factored and inverted as uhmk_chol_diag, the rows below
 * solved as uhmk_chol_trsm and ABR updated as uhmk_gemm_update(schur = 1,
 * transb = 1, tile = 64, mma) -- the same operations in the same order, so
 * the same bits as those four launches, in one launch per tree level.
 * Replaces op_merge + chol_hier per element, leaves first
 * (UHM/src/operation/element.cxx:200-241,
 * UHM/src/matrix/uhm/fla/chol/decompose.cxx:85-98).  prec: storage of the
 * fronts; mma: 0 FP64 products (FP64 fronts), 1 FP32 products. */
int uhmk_chol_front_small(const uhmk_front* fronts, const int* flist, int n_fronts, const int* src_beg,
                          const uhmk_source* srcs, const int* inv, double* arena64, float* arena32, int prec, int mma,
                          const double* stage, int* info, hipStream_t stream);

/* Low-precision Schur complement products (the final ABR -= ABL ATR /
 * ABL ABL^T of lu_piv_hier / chol_hier,
 * UHM/src/matrix/uhm/fla/lu_piv/decompose.cxx:148-149,
 * UHM/src/matrix/uhm/fla/chol/decompose.cxx:95-96) on packed BF16 / FP16
 * operands with FP32 accumulation.  Packed layout of front f at element
 * off[f] of lp: A = ABL as ss_pad rows of ldk entries (K contiguous), for LU
 * B = ATR^T after it; ldk = fs rounded up to 64, ss_pad = ss rounded up to
 * 128, zero padding; uhmk_schur_lp_size gives the entries a front needs.
 * FP16 operands carry a power-of-two scale per front and operand, chosen
 * from amax (2 unsigned per front: the float bits of max |entry| of A and
 * B, zero before uhmk_schur_lp_pack) so that the largest entry lies in
 * [2^14, 2^15); BF16 ones are unscaled.  Rounding: nearest even, once,
 * from the storage value.
 *   uhmk_schur_lp_pack  items {f, row tile, k tile, operand (0 A, 1 B)}
 *                       over 64 x 64 tiles of the padded packed operands
 *   uhmk_schur_lp       C -= A B^T (Cholesky: B = A) on tile x tile blocks
 *                       of ABR (tile 64 or 128; 64 is the faster one, eight
 *                       workgroups per CU), items {f, i0, j0, 0} (front
 *                       indices, tiles cut at n); C updated in its storage
 *                       type */
long long uhmk_schur_lp_size(int n, int fs, int lu);
int uhmk_schur_lp_pack(const uhmk_front* fronts, const uhmk_item* items, int n_items, int fp16, const void* arena,
                       int prec, const long long* off, unsigned* amax, unsigned short* lp, hipStream_t stream);
int uhmk_schur_lp(const uhmk_front* fronts, const uhmk_item* items, int n_items, int lu, int tile, int fp16,
                  void* arena, int prec, const long long* off, const unsigned* amax, const unsigned short* lp,
                  hipStream_t stream);

/* Cholesky updates of FP64 fronts with FP32 MFMA on an FP32 copy of their
 * operand: trailing updates C -= L(:, [c0, c1)) L(:, [c0, c1))^T on the rows
 * [c1, n), c1 = min(c0 + nb, fs) (LINAL/src/dense/chol.cxx:86-90), or with
 * schur != 0 the final Schur complement ABR -= ABL ABL^T (c0 = 0, c1 = fs;
 * UHM/src/matrix/uhm/fla/chol/decompose.cxx:95-96).  The operand, rows
 * [c1, n) of the columns [c0, c1), rounded to FP32 (nearest even, once)
 * into a column-major copy at element 64 * item.c of P (leading dimension
 * n - c1 rounded up to 4, padding rows zero); uhmk_p32_size gives the
 * entries a front needs.  The products read the copy with 16-byte loads; C
 * (the FP64 front) is the same bits as uhmk_gemm_update's NT update with
 * mma = 1 on the front itself.  prec must be 0 (-1006 otherwise).
 *   uhmk_p32_pack         items {f, 64-row tile of the copy, 64-column
 *                         tile, offset / 64}
 *   uhmk_gemm_update_p32  items {f, i0, j0, offset / 64}: 64 x 64 tiles of
 *                         C (front indices, lower ones only) */
long long uhmk_p32_size(int n, int fs, int c0, int nb, int schur);
int uhmk_p32_pack(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, int schur,
                  const void* arena, int prec, float* P, hipStream_t stream);
int uhmk_gemm_update_p32(const uhmk_front* fronts, const uhmk_item* items, int n_items, int c0, int nb, int schur,
                         void* arena, int prec, const float* P, hipStream_t stream);

/* ---- solve ---------------------------------------------------------------- */

/* Composed row permutation of each LU front (perm[i] = row of the
 * unpivoted right-hand side that the relative pivots bring to position i,
 * swaps k <-> k + p[k] in order k = 0 .. fs-1, as FLA_Apply_pivots does in
 * solve_lu_piv_1_hier, UHM/src/matrix/uhm/fla/lu_piv/solve.cxx:161-187).
 * Computed once per factorization; perm has the layout of piv. */
int uhmk_lu_perm(const uhmk_front* fronts, const int* flist, int n_fronts, const int* piv, int* perm,
                 hipStream_t stream);

/* Forward-solve gather for one tree level: x_f = b_f + sum of the
 * children's xb in merge order (Helper_::_merge_rhs, helper.hxx:253-301),
 * the factor rows read through perm for LU.  items: {f, 0, first row, 0},
 * 256 rows each; all n_rhs columns. */
int uhmk_solve_gather(const uhmk_front* fronts, const uhmk_item* items, int n_items, int lu, const int* src_beg,
                      const uhmk_source* srcs, const int* inv, const int* perm, double* x, const double* b, int nrhs,
                      hipStream_t stream);

/* Triangular solve of the diagonal super-block [sb * sb_size,
 * min((sb + 1) * sb_size, fs)) of ATL for every front in flist (one
 * workgroup per front; fronts with fs <= sb * sb_size are skipped;
 * sb_size a multiple of 64, at most 1024): fwd = 1: x_t <- L^-1 x_t (unit
 * lower for LU); fwd = 0: x_t <- U^-1 x_t (LU) or L^-T x_t (Cholesky).
 * chol: 0 LU, 1 Cholesky, 2 Cholesky whose 64 x 64 diagonal blocks hold
 * the inverse written by uhmk_chol_diag (applied as a GEMV per block).
 * With sb_size >= fs this is the whole ATL solve; otherwise the couplings
 * between super-blocks go through uhmk_solve_gemv.  arena holds the fronts
 * in FP64 (prec 0) or FP32 (prec 1); arithmetic is FP64.  Replaces the
 * triangular parts of solve_{lu_piv,chol}_{1,2}_hier
 * (lu_piv/solve.cxx:161-231, chol/solve.cxx:152-202). */
int uhmk_solve_trsv(const uhmk_front* fronts, const int* flist, int n_fronts, int max_fs, int chol, int fwd,
                    int sb, int sb_size, const void* arena, int prec, double* x, int nrhs, hipStream_t stream);

/* Off-diagonal solve updates.  cols = 0: x[rows] -= F[rows, a:c) x[a:c)
 * over row chunks, items {f, a, first row, c}, rows [b, min(b + RT,
 * b < fs ? fs : n)), RT = 16 (FP64) or 32 (FP32): xb -= ABL x_t, x_t -=
 * ATR xb and super-block couplings (dst must be x or null).  cols = 1:
 * dst[col] -= sum over rows [a, c) of F(row, col) x[row] for columns
 * [b, min(b + 16, fs)), items {f, a, first column, c}: x_t -= ABL^T xb and
 * the Cholesky backward couplings (dst: x when null). */
int uhmk_solve_gemv(const uhmk_front* fronts, const uhmk_item* items, int n_items, int cols, const void* arena,
                    int prec, double* x, double* dst, int nrhs, hipStream_t stream);

/* One 64-row block step k0 of the Cholesky solves of a tree level, for
 * factors whose diagonal blocks carry the inverse from uhmk_chol_diag
 * (solve_chol_{1,2}_hier, UHM/src/matrix/uhm/fla/chol/solve.cxx:152-202).
 * fwd = 1: y = W x[k0:k0+64) goes to z, and x[rows] -= L[rows, block] y for
 * the 64-row chunk of the item (items {f, 0, first row, 0}: k0 for the
 * block itself, else k0 + kb + 64 m with kb = min(64, fs - k0); rows run
 * into ABL, so the xb update is included).  fwd = 0: y = W^T z[k0:k0+64) goes to x, and
 * z[cols] -= L[block, cols]^T y for the 64-column chunk of the item (items
 * {f, 0, first column <= k0, 0}).  x and z have the front layout of the
 * right-hand sides. */
int uhmk_solve_chol_step(const uhmk_front* fronts, const uhmk_item* items, int n_items, int fwd, int k0,
                         const void* arena, int prec, double* x, double* z, int nrhs, hipStream_t stream);

/* The Cholesky solves of one tree level in one launch per direction, for
 * factors whose diagonal blocks carry the inverse from uhmk_chol_diag /
 * uhmk_chol_group (solve_chol_{1,2}_hier,
 * UHM/src/matrix/uhm/fla/chol/solve.cxx:152-202): the 64-row block steps of
 * uhmk_solve_chol_step and the ABL products of uhmk_solve_gemv, with the
 * dependencies between the blocks resolved inside the launch.  fwd = 1:
 * items {f, 0, first row r0, 0} for every 64-row block of ATL (each front's
 * in increasing r0), then {f, 1, first row, 0} for every 64-row chunk of
 * ABL; z[0:fs) <- L^-1 x[0:fs) (z marked by uhmk_solve_mark beforehand),
 * x[fs:n) -= ABL z[0:fs), and the ATL rows of x are marked for the
 * backward launch.  fwd = 0: items {f, 0, r0, 0} for every block of ATL,
 * each front's in decreasing r0; x[0:fs) <- L^-T (z[0:fs) - ABL^T x[fs:n)).
 * x, z: the front layout of the right-hand sides; nrhs <= 8.  err: one int,
 * zero before; raised if a wait gave up (the results are then invalid). */
int uhmk_solve_chol_level(const uhmk_front* fronts, const uhmk_item* items, int n_items, int fwd,
                          const void* arena, int prec, double* x, double* z, int nrhs, int* err, hipStream_t stream);
/* p[0..n) = the mark of blocks not yet solved (a NaN bit pattern), as a
 * kernel (usable inside a captured graph). */
int uhmk_solve_mark(double* p, long long n, hipStream_t stream);

/* Block solves with many right-hand sides (FP64 MFMA over tiles of 64
 * right-hand sides; items {f, first right-hand side of the tile, first
 * row, 0}).  One 64-row block step k0 of the ATL triangular solves of a
 * tree level (the triangular parts of solve_{lu_piv,chol}_{1,2}_hier,
 * UHM/src/matrix/uhm/fla/lu_piv/solve.cxx:161-231,
 * UHM/src/matrix/uhm/fla/chol/solve.cxx:152-202): item row k0 is the
 * block itself, y = T_kk^-1 x_k goes to z (fwd = 1) or y = T_kk^-1 z_k to
 * x (fwd = 0); any other item row r is a 64-row chunk of ATL updated with
 * y: forward x_r -= F(r, k) y (r > k0), backward z_r -= F(r, k) y (LU) or
 * F(k, r)^T y (Cholesky) (r < k0).  mode 0: LU (unit lower L forward, U
 * backward), 1: Cholesky (L, L^T), 2: Cholesky whose diagonal blocks carry
 * W = L_kk^-1 from uhmk_chol_diag. */
int uhmk_solve_block_step(const uhmk_front* fronts, const uhmk_item* items, int n_items, int mode, int fwd, int k0,
                          const void* arena, int prec, double* x, double* z, int nrhs, hipStream_t stream);

/* The LU (mode 0) block step at k0 = 0 for items whose front has fs <= 32
 * (each the front's own block): one wavefront per item, the right-hand-side
 * columns in registers; the same bits as uhmk_solve_block_step. */
int uhmk_solve_block_step_small(const uhmk_front* fronts, const uhmk_item* items, int n_items, int fwd,
                                const void* arena, int prec, double* x, double* z, int nrhs, hipStream_t stream);

/* Level-batched off-diagonal products of the block solves, 64 rows x 64
 * right-hand sides per item: dir 0: x[ABL rows] -= ABL z[0:fs) (forward
 * xb update); dir 1: z[ATL rows] -= ATR x[fs:n) (LU backward); dir 2:
 * z[ATL rows] -= ABL^T x[fs:n) (Cholesky backward).  ksplit > 1 cuts each
 * item's K range into ksplit slices (more workgroups for levels with few
 * long products); the partial tiles (n_items * ksplit * 4096 doubles in
 * part) are summed in slice order by a second kernel.  The gemm / gemv
 * calls of solve_{lu_piv,chol}_{1,2}_hier on n_rhs columns. */
int uhmk_solve_block_gemm(const uhmk_front* fronts, const uhmk_item* items, int n_items, int dir,
                          const void* arena, int prec, double* x, double* z, int nrhs, int ksplit, double* part,
                          hipStream_t stream);

/* uhmk_solve_gather and uhmk_solve_branch for many right-hand sides: the
 * same results, bit for bit, with items of 64 rows ({f, 0, first row, 0})
 * and one wavefront per item and group of 8 columns. */
int uhmk_solve_gather_cols(const uhmk_front* fronts, const uhmk_item* items, int n_items, int lu,
                           const int* src_beg, const uhmk_source* srcs, const int* inv, const int* perm, double* x,
                           const double* b, int nrhs, hipStream_t stream);
int uhmk_solve_branch_cols(const uhmk_front* fronts, const uhmk_item* items, int n_items, const int* emap, double* x,
                           int nrhs, hipStream_t stream);

/* Backward branch: xb <- the parent's solution rows through emap
 * (Helper_::_branch_rhs, helper.hxx:302-331).  items {f, 0, first row, 0},
 * 256 rows each. */
int uhmk_solve_branch(const uhmk_front* fronts, const uhmk_item* items, int n_items, const int* emap, double* x,
                      int nrhs, hipStream_t stream);

/* p[0..n) = value, as a kernel (usable inside a captured graph, where
 * memset nodes are avoided). */
int uhmk_fill_int(int* p, long long n, int value, hipStream_t stream);

/* ---- inputs and residuals -------------------------------------------------- */

/* Synthetic element matrices K_e = s_e (M^T M + alpha I + beta S) and
 * b_e ~ N(0,1) from a Philox-4x32-10 stream, written into the leaf stages.
 * Stands in for Mesh_::random_matrix / random_spd_matrix
 * (UHM/src/mesh/matrix.cxx:293-309).  items: {f, element id, lmap offset, m}. */
int uhmk_generate_elements(const uhmk_front* fronts, const uhmk_item* items, int n_items, int max_m,
                           const int* lmap, unsigned long long seed, double alpha, double beta,
                           int scale_kind, double p0, double p1, double* stage, double* b, int nrhs,
                           int with_lhs, int with_rhs, hipStream_t stream);

/* CSR values of the assembled matrix from the leaf stages (DSparse_
 * assembly, UHM/src/interf/sparse.cxx:196-267). */
int uhmk_csr_values(long long nnz, const long long* cptr, const long long* cidx, const double* stage, double* val,
                    hipStream_t stream);

/* Assembled right-hand side column from the leaf right-hand sides. */
int uhmk_gather_rhs(long long nrow, const long long* bptr, const long long* bidx, const int* bn, const double* b,
                    double* bg, int col, hipStream_t stream);

/* Residual r = b - A x for all right-hand sides from the leaf input
 * matrices (A = sum_e P_e K_e P_e^T, the assembly of DSparse_::import_element,
 * UHM/src/interf/sparse.cxx:196-267), replacing check_* / get_residual
 * (UHM/src/mesh/lu_piv.cxx:184-196, UHM/src/mesh/matrix.cxx:241-275) for
 * many columns at once.  Compact leaf layout: leaf f's n x n_rhs block at
 * loff_f * n_rhs, leading dimension n.
 *   uhmk_leaf_gather    xl_e = P_e^T xg; items {f, 0, first row, loff_f},
 *                       lg[loff_f + i] = global row of leaf row i
 *   uhmk_leaf_mul       yl_e = K_e xl_e (FP64 MFMA); items {f, first
 *                       right-hand side of a 64-column tile, first row,
 *                       loff_f}; K_e from the leaf stages
 *   uhmk_leaf_residual  r(g) = b(g) - sum of the yl entries of row g (lpos:
 *                       compact position of each entry listed in bptr, bn
 *                       its leaf's n), leaf order; out[4 c + {0, 2, 3}] =
 *                       max over rows of |r|, |x|, |b| of column c (atomic
 *                       maxima: out zeroed by the caller).  With bg and out
 *                       null: the product r = A x (GMRES-IR operator)
 *   uhmk_rowsum_max     out = max row sum |A| of the CSR matrix (atomic
 *                       maximum: out zeroed by the caller)
 * xg, bg, r: ndof x n_rhs, column major. */
int uhmk_leaf_gather(const uhmk_front* fronts, const uhmk_item* items, int n_items, const long long* lg,
                     const double* xg, long long nglob, double* xl, int nrhs, hipStream_t stream);
int uhmk_leaf_mul(const uhmk_front* fronts, const uhmk_item* items, int n_items, const double* stage,
                  const double* xl, double* yl, int nrhs, hipStream_t stream);
int uhmk_leaf_residual(long long nrow, const long long* bptr, const long long* lpos, const int* bn,
                       const double* yl, const double* xg, const double* bg, double* r, int nrhs, double* out,
                       hipStream_t stream);
int uhmk_rowsum_max(long long nrow, const long long* rowptr, const double* val, double* out, hipStream_t stream);

/* Between per-front solution blocks and a global vector over the factor
 * rows, all n_rhs columns: dir 0 g = xf (fronts -> global), dir 1 xf = g
 * (global -> fronts), dir 2 g += xf (refinement update); the analogue of
 * copy_out for the whole mesh (UHM/src/mesh/interf.cxx:183-199). */
int uhmk_front_global_all(const uhmk_front* fronts, int n_fronts, const long long* goff, double* xf, double* g,
                          int nrhs, long long nglob, int dir, hipStream_t stream);

/* ---- GMRES-IR vector kernels ------------------------------------------------
 * Global vectors of ndof x n_rhs (column major); a basis V of k such blocks
 * at stride vstride.  Fixed-order reductions (bitwise reproducible).
 * Refinement slot: Mesh_::improve_chol / improve_lu_piv
 * (UHM/src/mesh/chol.cxx:136-149, UHM/src/mesh/lu_piv.cxx:232-245).
 *   uhmk_block_dots     out[i * n_rhs + c] = V_i(:, c) . w(:, c), i < k <= 32;
 *                       partial: uhmk_block_dots_scratch doubles
 *   uhmk_basis_combine  w = beta w + alpha sum_i coef[i * n_rhs + c] V_i(:, c)
 *   uhmk_scale_columns  w(:, c) *= s[c] */
long long uhmk_block_dots_scratch(long long nrow, int nrhs);
int uhmk_block_dots(const double* V, long long vstride, int k, const double* w, long long nrow, int nrhs,
                    double* partial, double* out, hipStream_t stream);
/* uhmk_block_dots without the caller's workspace: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops): it calls no hipMalloc / hipFree
 * and keeps its partial sums in a fixed device array of
 * uhmk_block_dots_noalloc_max() doubles (4 MiB; one copy per GPU, held until
 * the process ends).  Same results as uhmk_block_dots, bit for bit.  Returns
 * -1013 when uhmk_block_dots_scratch(nrow, nrhs) exceeds that maximum,
 * -1012 for k > 32 or nrhs > 65535.  Calls that overlap on one GPU (two
 * streams, two host threads) share the array and race. */
long long uhmk_block_dots_noalloc_max(void);
int uhmk_block_dots_noalloc(const double* V, long long vstride, int k, const double* w, long long nrow, int nrhs,
                            double* out, hipStream_t stream);
int uhmk_basis_combine(const double* V, long long vstride, int k, const double* coef, double alpha, double beta,
                       double* w, long long nrow, int nrhs, hipStream_t stream);
int uhmk_scale_columns(double* w, long long nrow, int nrhs, const double* s, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
