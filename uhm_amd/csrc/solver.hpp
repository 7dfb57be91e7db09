// Device-resident multifrontal factorization and solve.
//
// FrontSolver owns every front of a locked mesh in one HBM arena (each front
// a column-major n x n block, n = fs + ss, quadrants ATL/ABL/ATR/ABR), the
// pivots, the right-hand sides, the extend-add index maps and the per-level
// work lists.  The symbolic part (maps, offsets, work items) is built once
// per sparsity pattern; factor() then replays a fixed list of batched
// launches, one group per tree level, leaves first.
//
// It stands in for the reference's per-element Matrix_FLA_ objects
// (UHM/include/uhm/matrix/uhm/fla.hxx:65-195) driven by Scheduler_
// (UHM/src/operation/scheduler.cxx:210-328): the scheduler's generations are
// the batching unit here.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "mesh.hpp"
#include "uhmk.h"

namespace uhm {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() {}
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release();
  void alloc(size_t b);    // exactly b bytes (old contents dropped)
  void reserve(size_t b);  // at least b bytes; keeps a large enough buffer
  template <class T>
  T* as() const { return static_cast<T*>(p); }
  template <class T>
  void upload(const std::vector<T>& v, hipStream_t s);
};

// One launch of the level schedule.
struct Launch {
  // GEMM_DIAG: a Cholesky panel update whose diagonal tiles then factor the
  // next panel's diagonal block (items with c == 1)
  // CHOL_GROUP: every panel of a Cholesky panel group (diagonal blocks,
  // TRSM, in-group updates) in one persistent launch
  // LP_PACK: the Schur operands of a level packed into BF16 / FP16 for the
  // low-precision SCHUR launch that follows (precision modes 2 and 3)
  // CHOL_SMALL: assembly and whole Cholesky factorization (diagonal block,
  // TRSM, Schur complement) of every front of a level with fs <= 64, a
  // workgroup per front (front list)
  // SCHUR_LP: the Schur complement products of a level on the packed BF16 /
  // FP16 operands of LP_PACK
  // P32_PACK: the operand of the SCHUR_P32 / GEMM_P32 launch that follows
  // (Cholesky, FP64 fronts, FP32 products) copied to FP32 (items carry the
  // copy's offset / 64 in c)
  enum Kind {
    ASSEMBLE, LU_PANEL, LU_ROWS, CHOL_DIAG, CHOL_TRSM, GEMM, SCHUR, LU_FRONT, GEMM_DIAG, CHOL_GROUP, LP_PACK,
    CHOL_SMALL, SCHUR_LP, P32_PACK, SCHUR_P32, GEMM_P32
  } kind;
  int c0 = 0, nb = 0;
  long long off = 0;  // offset into the item (or front-list) array
  int count = 0;
  int extra = 0;      // max pivot rows (panel), transb (gemm / schur), schur (p32 pack)
  int level = 0;
  int tile = 64;      // gemm / schur tile edge
  int prec = 0;       // storage of the level's fronts: 0 FP64, 1 FP32
  int mma = 0;        // gemm / schur products: 0 FP64, 1 FP32, 2 BF16 operands (FP32 sums)
  double flops = 0;   // algorithmic flops of the launch
  double bytes = 0;   // algorithmic HBM bytes of the launch
  int stream = 0;     // 0: main stream, 1: lookahead stream
  int wait_ev = -1;   // event the launch's stream waits for first (-1: none)
  int record_ev = -1; // event recorded on the launch's stream after it (-1: none)
};
constexpr int N_LAUNCH_KINDS = 16;

// Size classes of fronts for the schedule's knobs: fs <= 64 (one panel),
// fs <= 512 (a few panels), larger.
constexpr int N_FRONT_CLASSES = 3;
inline int front_class(int fs) { return fs <= 64 ? 0 : (fs <= 512 ? 1 : 2); }

// One launch of the level solves.
struct SolveStep {
  // LEVEL: the Cholesky solves of a level in one launch (fwd as for CSTEP);
  // MARK: z marked as not yet solved and the error word cleared
  // BSMALL: the LU block step of fronts with fs <= 32 (one wavefront per item)
  enum Kind { GATHER, TRSV, GEMV, BRANCH, CSTEP, BSTEP, BGEMM, LEVEL, MARK, BSMALL } kind = GATHER;
  int off = 0, cnt = 0;            // items (gather / gemv / branch) or front list (trsv)
  int max_fs = 0, sb = 0, sb_size = 0;  // trsv
  int cols = 0;                    // gemv: row chunks (0) or column chunks (1); gather / branch:
                                   // 256-row items (0) or 64-row items x column groups (1)
  int prec = 0;                    // storage of the level's fronts
  int fwd = 0;                     // trsv / cstep: forward (1) or backward (0)
  int k0 = 0;                      // cstep / bstep: first row of the block
  int mode = 0;                    // bstep: 0 LU, 1 Cholesky, 2 Cholesky with W; bgemm: direction
  int ks = 1;                      // bgemm: K slices (split-K for levels with few, long products)
};

// Per-kind totals of a profiled factorization (device events per launch).
struct KernelStats {
  double ms[N_LAUNCH_KINDS] = {0};
  double flops[N_LAUNCH_KINDS] = {0};
  double bytes[N_LAUNCH_KINDS] = {0};
  int launches[N_LAUNCH_KINDS] = {0};
};

// Device bytes held by the library's buffers now and at most so far.
long long device_bytes_used();
long long device_bytes_max();

class FrontSolver {
 public:
  FrontSolver(Mesh_* m, int n_rhs);
  ~FrontSolver();

  void create_buffers();
  bool buffers_created() const { return created_; }

  // copy-in staging (host, front order, accumulating)
  void stage_lhs(int front, int m, const int* map, const double* K);
  void stage_rhs(int front, int m, int nrhs, const int* map, const double* b);
  void set_rhs();

  // synthetic element matrices / rhs generated on the device
  void generate(unsigned long long seed, double alpha, double beta, int scale_kind, double p0, double p1,
                int with_lhs, int with_rhs);

  int factor(int method);  // 0, or element id + 1 of the first zero-pivot / non-SPD front

  // Factorization reuse (Mesh_::check_reuse, mesh.cxx:159-225, and the skip
  // in op_decompose, operation/element.cxx:204-206): fronts of elements
  // flagged reusable take their factored entries (and LU pivots) from the
  // solver `prev` of the previous pattern at the next factorization, and
  // every factorization skips them.  prev must have factored; a reusable
  // front whose size differs there is refactored with its ancestors.
  void set_reuse_source(std::unique_ptr<FrontSolver> prev);
  int n_reused() const;
  bool factored() const { return factored_; }
  void solve(int method);
  void check(int method);
  double get_residual() const { return residual_1norm_; }
  double backward_error() const { return eta_; }
  const double* residual_norms() const { return norms_; }

  // results
  void fetch_x(int front, int m, int nrhs, const int* map, double* out);
  void fetch_front(int front, int m, const int* map, double* out);
  void fetch_front_raw(int front, double* out);  // n x n, column major
  void fetch_pivots(int front, int* out);
  void fetch_stage(int front, double* out);      // n x n stage (assembled element input)
  void fetch_rhs_front(int front, double* out);  // n x n_rhs right-hand side block
  void fetch_x_global(double* out);              // ndof x nrhs
  void fetch_b_global(double* out);

  // symbolic views for tests / tools
  int n_fronts() const { return (int)fr_.size(); }
  const uhmk_front& front(int f) const { return fr_[f]; }
  Element_* element_of(int f) const { return order_[f]; }
  int level_of(int f) const { return front_level_[f]; }
  const std::vector<int>& emap() const { return emap_; }
  int n_levels() const { return (int)levels_.size(); }
  long long ndof() const { return ndof_; }
  int n_rhs() const { return nrhs_; }
  void csr(std::vector<long long>& rowptr, std::vector<int>& col, std::vector<double>& val);

  // Precision of the fronts.  mode 0: every front FP64 (FP64 MFMA).
  // mode 1 (mixed): fronts of the top `top_levels` tree levels stay FP64
  // (their pivot blocks factorize in FP64), all other fronts are stored in
  // FP32, and every Schur / trailing update runs as FP32 MFMA products.
  // mode 2 / 3: as mode 1 with the final Schur complement products on BF16 /
  // FP16 operands (FP32 accumulation), the operands packed once per front.
  // Changing it re-lays out the arenas and needs a new factorization.
  void set_precision(int mode, int top_levels);
  int precision_mode() const { return prec_mode_; }
  int level_precision(int L) const { return level_prec_[L]; }

  // Iterative refinement on the last factorization: x <- x + M^-1 (b - A x)
  // with the residual in FP64 against the assembled matrix, until the
  // normwise backward error of every column is <= tol or max_iter
  // corrections were applied.  Returns the number of corrections; the
  // backward error after each step is in refine_history().
  int refine(int method, double tol, int max_iter);
  const std::vector<double>& refine_history() const { return refine_hist_; }
  // Correction kind: 0 classical (d = M^-1 r), 1 GMRES-IR (d from GMRES(m)
  // on M^-1 A d = M^-1 r, stopped at inner_tol relative to ||M^-1 r||),
  // 2 classical while each step cuts the backward error at least tenfold,
  // GMRES-IR from the first step that does not.
  void set_refinement(int kind, int restart, double inner_tol);
  int refine_kind() const { return refine_kind_; }
  int refine_solves() const { return refine_solves_; }  // M^-1 applications of the last refinement
  double last_refine_ms() const { return t_refine_ms_; }

  // tuning / timing
  void set_block(int nb) {
    if (nb != nb_) schedule_method_ = 0;
    nb_ = nb;
  }
  // Schedule knobs: key 1 = Cholesky panels per trailing update, key 2 =
  // mean tile extent from which a GEMM launch uses 128 x 128 tiles, key 3 =
  // number of right-hand sides from which the solves run as block solves
  // (MFMA products over 64-column tiles of right-hand sides), key 4 =
  // largest LU front (n <= 64) factored by the whole-front kernel (0: none),
  // key 5 = Cholesky panel lookahead (1 on, 0 off).
  void set_tuning(int key, int value);
  int get_tuning(int key) const;
  // Times factorizations of the current values over the knobs and keeps the
  // fastest setting; returns its factorization time in ms.
  double autotune(int method);
  // Cholesky panels per trailing update (1: update after every panel)
  void set_chol_group(int g) {
    if (g < 0) g = 0;
    if (g != chol_group_) schedule_method_ = 0;
    chol_group_ = g;
  }
  int chol_group() const { return chol_group_ ? chol_group_ : 6; }
  // per size class: panel group, GEMM tile threshold, group-launch policy
  // (0 / 0 / -1: the global knob)
  int class_group(int c) const { return cls_group_[c] > 0 ? cls_group_[c] : chol_group(); }
  void set_profile(bool on) { profile_ = on; }
  // Graph mode: the factorization's launch sequence (and each solve's) is
  // captured once per pattern / method / precision and replayed, for
  // re-factorization of new values on a fixed sparsity pattern.  Profiled
  // factorizations always run eagerly.
  void set_graph(bool on);
  bool graph_mode() const { return use_graph_; }
  const KernelStats& kernel_stats() const { return stats_; }
  int block() const { return nb_; }
  double last_factor_ms() const { return t_factor_ms_; }
  double last_solve_ms() const { return t_solve_ms_; }
  size_t device_bytes() const;
  hipStream_t stream() const { return stream_; }

 private:
  void build_symbolic();
  void build_schedule(int method);
  void build_csr();
  void build_solve_schedule(int method);
  void build_block_solve_schedule(int method);
  bool solve_schedule_current(int method) const;
  void layout_arenas();
  void* arena_ptr(int prec) const { return prec ? d_arena32_.p : d_arena_.p; }
  void residual_all(std::vector<double>& h);  // 4 per column: max|r|, max row sum |A|, max|x|, max|b|
  void apply_matrix(const double* v, double* out);             // out = A v (global vectors)
  void apply_factors(bool chol, const double* r, double* out);  // M^-1 r into d_x_ (fronts) and out (global, if any)
  void gmres_correction(bool chol);                             // x += GMRES(m) correction from r
  void run_solve_steps(const std::vector<SolveStep>& steps, bool chol, const double* rhs);
  void run_solve(bool chol, const double* rhs);
  void enqueue_factor(int method, std::vector<hipEvent_t>* pev);
  hipGraphExec_t capture(const std::function<void()>& body);
  void drop_graphs();
  void refresh_csr();
  void upload_stage();
  void upload_rhs();
  void check_hip(hipError_t e, const char* what) const;
  void check_k(int rc, const char* what) const;

  Mesh_* mesh_;
  int nrhs_;
  int nb_ = 64;
  int chol_group_ = 0;  // 0: 6 panels (measured best on the FP64, FP32 and low-precision paths)
  int tile_threshold_ = 8192;  // mean tile extent from which GEMM launches use 128 x 128 tiles
  int block_rhs_ = 8;
  int front_small_ = 64;  // LU fronts with n <= this (at most 64) factor in one whole-front launch
  int lookahead_ = 0;     // Cholesky: next panel's diagonal block and TRSM on the lookahead stream
  int fuse_diag_ = 1;     // Cholesky: next panel's diagonal block factored by the update's diagonal tile
  int small_fused_ = 0;  // Cholesky levels of at least this many fronts, all with fs <= 64, run
                         // as one whole-front launch (uhmk_chol_front_small); 0 never (default:
                         // the four level launches measured faster, see set_tuning key 9)
  int cls_group_[N_FRONT_CLASSES] = {0, 0, 0};
  int cls_tile_[N_FRONT_CLASSES] = {0, 0, 0};
  int cls_gkern_[N_FRONT_CLASSES] = {-1, -1, -1};
  int lp_min_ss_ = 768;   // precision modes 2 / 3: levels whose Schur blocks average fewer rows take
                          // FP32 products instead of packed BF16 / FP16 ones
  int lu_panel_ll_ = 1;   // LU panels: left-looking register-block kernel (1) or the right-looking
                          // one (0); the same bits on FP64 fronts
  int lu_lookahead_ = 1;  // LU: the next panel on the lookahead stream beside the rest of the update
  int svb_small_ = 1;     // LU block solves: fronts with fs <= 32 by one wavefront per item
  int schur_p32_ = 1;     // Cholesky, FP32 products on FP64 fronts: the Schur and trailing-update products
                          // on an FP32 copy of their operand (16-byte loads) instead of element loads
  int lu_rows_w_ = 32;    // LU row kernel: columns (ABL rows) per workgroup (measured best of 16, 32, 64)
  int solve_level_ = 1;   // Cholesky solves with inverses: one launch per level and direction (1) or
                          // one per 64-row block step (0)
  int group_kernel_ = 2;  // Cholesky: panel groups in one persistent launch (uhmk_chol_group): 0 never,
                          // 1 every level, 2 levels with at most 16 fronts
  int flag_slots_ = 1;    // ints of d_flags_ the schedule uses (progress flags of the group launches)
  hipStream_t side_stream_ = nullptr;
  std::vector<hipEvent_t> la_ev_;
  int final_wait_ev_ = -1;
  bool created_ = false;
  hipStream_t stream_ = nullptr;

  std::vector<Element_*> order_;
  std::vector<uhmk_front> fr_;
  std::vector<int> front_level_;
  std::vector<std::vector<int>> levels_;
  std::vector<int> emap_;          // concatenated child -> parent maps
  std::vector<int> inv_;           // concatenated parent -> child inverse maps
  std::vector<int> src_beg_;
  std::vector<uhmk_source> srcs_;
  std::vector<int> lmap_;          // leaf: element-local (user) order -> front index
  std::vector<int> lmap_off_;
  std::vector<long long> goff_;    // global offset of each front's factor block
  long long ndof_ = 0;
  long long arena_elems_ = 0, arena32_elems_ = 0, x_elems_ = 0, stage_elems_ = 0;
  int prec_mode_ = 0, top_fp64_ = 2;
  std::vector<int> level_prec_;
  int piv_elems_ = 0;

  // host staging
  // copy-in batch: the element blocks of copy_in calls in call order (values
  // and front positions packed), scattered into the stages / right-hand
  // sides by one device launch at the next upload (uhmk_copy_in)
  std::vector<uhmk_cblock> cin_blocks_;
  std::vector<double> cin_vals_;
  std::vector<int> cin_maps_;
  std::vector<char> cin_touched_;  // per (front, side): accumulating since creation / the last generate
  DevBuf d_cin_vals_, d_cin_maps_, d_cin_blocks_, d_cin_items_;
  void flush_copy_in();
  bool stage_dirty_ = false, rhs_dirty_ = false;

  // schedule
  int schedule_method_ = 0;
  std::vector<uhmk_item> items_;
  std::vector<int> flists_;
  std::vector<Launch> launches_;
  std::vector<uhmk_item> gen_items_;
  std::vector<std::vector<int>> level_lists_off_;
  // solve schedule
  int solve_method_ = 0;
  std::vector<uhmk_item> sitems_;
  std::vector<int> sflist_;
  std::vector<SolveStep> sfwd_, sbwd_;
  int n_fsl_ = 0;

  // device buffers
  DevBuf dF_, d_emap_, d_inv_, d_srcbeg_, d_srcs_, d_lmap_, d_goff_;
  DevBuf d_items_, d_flists_, d_gen_items_, d_lvl_, d_sitems_, d_sflist_, d_fsl_, d_perm_, d_pperm_;
  // panel permutation buffer of an LU panel / row launch: one of two per
  // front, alternating by panel index (the lookahead panel writes the other)
  int* pperm_of(const Launch& l) {
    return d_pperm_.as<int>() + (size_t)((l.c0 / (l.nb > 0 ? l.nb : 1)) & 1) * fr_.size() * 257;
  }
  DevBuf d_arena_, d_arena32_, d_piv_, d_x_, d_b_, d_stage_, d_info_;
  DevBuf d_flags_;  // progress flags of the Cholesky group launches, [0]: waits that gave up
  DevBuf d_serr_;   // level solves: waits that gave up
  // low-precision Schur operands (modes 2, 3): packed buffer, per-front
  // element offsets, per-front operand maxima (FP16 scales)
  DevBuf d_lp_, d_lpoff_, d_lpamax_;
  DevBuf d_p32_;  // FP32 operand copies of the SCHUR_P32 / GEMM_P32 launches (one launch's at a time)
  long long lp_elems_ = 0;
  void check_solve_err();
  DevBuf d_rf_;  // refinement right-hand side (front layout)
  DevBuf d_z_;   // second solve vector (front layout) of the block-step Cholesky solves
  DevBuf d_part_;  // split-K partial tiles of the block-solve products
  // assembled matrix (CSR) for residuals
  bool csr_built_ = false, vals_dirty_ = true, bg_dirty_ = true;
  long long nnz_ = 0;
  DevBuf d_rowptr_, d_colidx_, d_val_, d_cptr_, d_cidx_, d_bptr_, d_bidx_, d_xg_, d_bg_, d_r_, d_norms_, d_bn_;
  // leaf-product residual: compact leaf layout, its items and buffers
  DevBuf d_lg_, d_lpos_, d_lgi_, d_lmi_, d_xl_, d_yl_, d_amax_;
  int n_lgi_ = 0, n_lmi_ = 0;
  // GMRES-IR
  int refine_kind_ = 0, gmres_m_ = 20, refine_solves_ = 0;
  double gmres_tol_ = 1e-6;
  DevBuf d_V_, d_dpart_, d_dout_, d_coef_;
  double amax_ = 0.0;  // max row sum |A| of the current values

  double residual_1norm_ = 0.0, eta_ = 0.0;
  double norms_[4] = {0, 0, 0, 0};
  double t_factor_ms_ = 0, t_solve_ms_ = 0, t_refine_ms_ = 0;
  std::vector<double> refine_hist_;
  int last_method_ = 0;
  int factor_nb_ = 64;  // panel width of the last factorization
  int solve_inv_ = -1;  // solve schedule built for inverse-carrying Cholesky factors (1) or not (0)
  int solve_block_ = -1;  // solve schedule built as block solves (1) or per right-hand side (0)
  bool factored_ = false;
  std::vector<char> reuse_;                 // per front: factored entries kept, not refactored
  std::unique_ptr<FrontSolver> reuse_src_;  // fronts of the previous pattern, until adopted
  std::vector<std::vector<int>> flv_;       // per level: the fronts a factorization processes
  void adopt_reused();
  bool profile_ = false;
  bool use_graph_ = false;
  hipGraphExec_t factor_graph_ = nullptr;
  int factor_graph_method_ = 0;
  std::map<const double*, hipGraphExec_t> solve_graphs_;
  KernelStats stats_;
};

}  // namespace uhm
