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
  void alloc(size_t b);
  template <class T>
  T* as() const { return static_cast<T*>(p); }
  template <class T>
  void upload(const std::vector<T>& v, hipStream_t s);
};

// One launch of the level schedule.
struct Launch {
  enum Kind { ASSEMBLE, LU_PANEL, LU_ROWS, CHOL_DIAG, CHOL_TRSM, GEMM, SCHUR } kind;
  int c0 = 0, nb = 0;
  long long off = 0;  // offset into the item (or front-list) array
  int count = 0;
  int extra = 0;      // max pivot rows (panel), transb (gemm / schur)
  int level = 0;
  int tile = 64;      // gemm / schur tile edge
  double flops = 0;   // algorithmic flops of the launch
  double bytes = 0;   // algorithmic HBM bytes of the launch
};
constexpr int N_LAUNCH_KINDS = 7;

// Per-kind totals of a profiled factorization (device events per launch).
struct KernelStats {
  double ms[N_LAUNCH_KINDS] = {0};
  double flops[N_LAUNCH_KINDS] = {0};
  double bytes[N_LAUNCH_KINDS] = {0};
  int launches[N_LAUNCH_KINDS] = {0};
};

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

  // tuning / timing
  void set_block(int nb) {
    if (nb != nb_) schedule_method_ = 0;
    nb_ = nb;
  }
  void set_profile(bool on) { profile_ = on; }
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
  void upload_stage();
  void upload_rhs();
  void check_hip(hipError_t e, const char* what) const;
  void check_k(int rc, const char* what) const;

  Mesh_* mesh_;
  int nrhs_;
  int nb_ = 64;
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
  long long arena_elems_ = 0, x_elems_ = 0, stage_elems_ = 0;
  int piv_elems_ = 0;

  // host staging
  std::vector<std::vector<double>> hstage_;
  std::vector<std::vector<double>> hrhs_;
  bool stage_dirty_ = false, rhs_dirty_ = false;

  // schedule
  int schedule_method_ = 0;
  std::vector<uhmk_item> items_;
  std::vector<int> flists_;
  std::vector<Launch> launches_;
  std::vector<uhmk_item> gen_items_;
  std::vector<std::vector<int>> level_lists_off_;

  // device buffers
  DevBuf dF_, d_emap_, d_inv_, d_srcbeg_, d_srcs_, d_lmap_, d_goff_;
  DevBuf d_items_, d_flists_, d_gen_items_, d_lvl_;
  DevBuf d_arena_, d_piv_, d_x_, d_b_, d_stage_, d_info_;
  // assembled matrix (CSR) for residuals
  bool csr_built_ = false;
  long long nnz_ = 0;
  DevBuf d_rowptr_, d_colidx_, d_val_, d_cptr_, d_cidx_, d_bptr_, d_bidx_, d_xg_, d_bg_, d_r_, d_norms_;

  double residual_1norm_ = 0.0, eta_ = 0.0;
  double norms_[4] = {0, 0, 0, 0};
  double t_factor_ms_ = 0, t_solve_ms_ = 0;
  int last_method_ = 0;
  bool factored_ = false;
  bool profile_ = false;
  KernelStats stats_;
};

}  // namespace uhm
