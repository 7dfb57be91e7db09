// FrontSolver: symbolic plan + level-by-level drivers for the GPU
// multifrontal factorization, solve and residual check.
#include "solver.hpp"

#include <chrono>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <stdexcept>
#include <unordered_map>


namespace uhm {

// ------------------------------------------------------------------ DevBuf
// Device bytes held by every DevBuf of the process and their high-water
// mark: the accounting of the reference's process-global buffer_used
// (UHM/src/matrix/uhm/common.cxx:42-44), kept with atomics.
static std::atomic<long long> g_dev_bytes{0}, g_dev_bytes_max{0};

long long device_bytes_used() { return g_dev_bytes.load(); }
long long device_bytes_max() { return g_dev_bytes_max.load(); }

void DevBuf::release() {
  if (p) {
    (void)hipFree(p);
    g_dev_bytes -= (long long)bytes;
  }
  p = nullptr;
  bytes = 0;
}

void DevBuf::alloc(size_t b) {
  release();
  if (b == 0) b = 256;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) {
    p = nullptr;
    throw std::runtime_error(std::string("hipMalloc failed (") + std::to_string(b) + " bytes): " +
                             hipGetErrorString(e));
  }
  bytes = b;
  const long long now = (g_dev_bytes += (long long)b);
  long long prev = g_dev_bytes_max.load();
  while (now > prev && !g_dev_bytes_max.compare_exchange_weak(prev, now)) {
  }
}

void DevBuf::reserve(size_t b) {
  if (!p || bytes < b) alloc(b);
}

template <class T>
void DevBuf::upload(const std::vector<T>& v, hipStream_t s) {
  alloc(v.size() * sizeof(T));
  if (!v.empty()) {
    hipError_t e = hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) throw std::runtime_error(std::string("upload: ") + hipGetErrorString(e));
  }
}

// ------------------------------------------------------------ FrontSolver
FrontSolver::FrontSolver(Mesh_* m, int n_rhs) : mesh_(m), nrhs_(n_rhs) { build_symbolic(); }

FrontSolver::~FrontSolver() {
  drop_graphs();
  for (hipEvent_t e : la_ev_) (void)hipEventDestroy(e);
  if (side_stream_) (void)hipStreamDestroy(side_stream_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void FrontSolver::check_hip(hipError_t e, const char* what) const {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void FrontSolver::check_k(int rc, const char* what) const {
  if (rc != 0) throw std::runtime_error(std::string(what) + ": kernel launch failed, code " + std::to_string(rc));
}

static inline long long align_up(long long v, long long a) { return (v + a - 1) / a * a; }

void FrontSolver::layout_arenas() {
  // Fronts of FP64 levels are packed into the FP64 arena, the others into
  // the FP32 arena; each offset is 512-byte aligned.
  const int nl = (int)levels_.size();
  level_prec_.assign(nl, 0);
  if (prec_mode_ != 0)
    for (int L = 0; L < nl - top_fp64_; ++L) level_prec_[L] = 1;
  long long a64 = 0, a32 = 0;
  for (int f = 0; f < (int)fr_.size(); ++f) {
    uhmk_front& F = fr_[f];
    F.prec = level_prec_[front_level_[f]];
    const long long sz = (long long)F.n * F.n;
    if (F.prec) {
      F.off = a32;
      a32 = align_up(a32 + sz, 128);
    } else {
      F.off = a64;
      a64 = align_up(a64 + sz, 64);
    }
  }
  arena_elems_ = a64;
  arena32_elems_ = a32;
}

void FrontSolver::set_precision(int mode, int top_levels) {
  if (mode < 0 || mode > 3)
    throw std::runtime_error("set_precision: mode must be 0 (FP64), 1 (mixed FP32), 2 (mixed, BF16 Schur) or 3 "
                             "(mixed, FP16 Schur)");
  if (top_levels < 0) throw std::runtime_error("set_precision: top_levels must be >= 0");
  if (mode == prec_mode_ && top_levels == top_fp64_) return;
  prec_mode_ = mode;
  top_fp64_ = top_levels;
  drop_graphs();
  layout_arenas();
  // fronts kept from a previous pattern are lost with a re-laid-out arena
  // once adopted; before adoption they are converted to the new storage
  if (!reuse_src_ && created_) reuse_.assign(fr_.size(), 0);
  schedule_method_ = 0;
  solve_method_ = 0;
  factored_ = false;
  if (created_) {
    check_hip(hipStreamSynchronize(stream_), "set_precision");
    dF_.upload(fr_, stream_);
    d_arena_.release();
    d_arena32_.release();
    d_arena_.alloc(std::max(1LL, arena_elems_) * sizeof(double));
    d_arena32_.alloc(std::max(1LL, arena32_elems_) * sizeof(float));
    check_hip(hipStreamSynchronize(stream_), "set_precision");
  }
}

void FrontSolver::build_symbolic() {
  // Front order = scheduler order: generations leaf-first, element id order
  // inside a generation (scheduler.cxx:157-178, 306-328).
  const auto& lv = mesh_->levels();
  order_.clear();
  levels_.assign(lv.size(), {});
  front_level_.clear();
  for (size_t L = 0; L < lv.size(); ++L)
    for (Element_* e : lv[L]) {
      e->front = (int)order_.size();
      levels_[L].push_back(e->front);
      front_level_.push_back((int)L);
      order_.push_back(e);
    }
  const int nf = (int)order_.size();
  fr_.assign(nf, uhmk_front{});
  long long xoff = 0, soff = 0, g = 0;
  int poff = 0;
  for (int f = 0; f < nf; ++f) {
    Element_* e = order_[f];
    uhmk_front& F = fr_[f];
    F.fs = e->fs;
    F.n = e->fs + e->ss;
    F.xoff = xoff;
    xoff = align_up(xoff + (long long)F.n * nrhs_, 8);
    F.poff = poff;
    poff += F.fs;
    F.parent = e->parent ? e->parent->front : -1;
    F.id = e->id;
    F.soff = -1;
    if (e->is_leaf()) {
      F.soff = soff;
      soff = align_up(soff + (long long)F.n * F.n, 64);
    }
    F.emap = -1;
  }
  layout_arenas();
  x_elems_ = xoff;
  stage_elems_ = soff;
  piv_elems_ = poff;
  // global offsets: factor blocks concatenate in numbering order
  goff_.assign(nf, 0);
  for (int f = 0; f < nf; ++f) {
    Element_* e = order_[f];
    if (e->fs) {
      goff_[f] = e->factor.front().first->offset;
      if (goff_[f] != g) throw std::runtime_error("build_symbolic: factor numbering is not contiguous in front order");
    } else {
      goff_[f] = g;
    }
    g += e->fs;
  }
  ndof_ = g;

  // child -> parent maps (unified parent index) and parent -> child inverses
  emap_.clear();
  inv_.clear();
  src_beg_.assign(nf + 1, 0);
  srcs_.clear();
  for (int f = 0; f < nf; ++f) {
    Element_* e = order_[f];
    src_beg_[f] = (int)srcs_.size();
    if (e->is_leaf()) srcs_.push_back(uhmk_source{-1, -1});
    for (Element_* c : e->children) {
      const int cf = c->front;
      std::vector<Mapper_> runs = build_mapper(e, c);
      fr_[cf].emap = (int)emap_.size();
      const int base = (int)emap_.size();
      emap_.resize(base + c->ss, -1);
      const int ib = (int)inv_.size();
      inv_.resize(ib + fr_[f].n, -1);
      for (auto& r : runs)
        for (int d = 0; d < r.n_dof; ++d) {
          const int pidx = (r.fs_p ? e->fs : 0) + r.offs_p + d;
          emap_[base + r.offs_c + d] = pidx;
          inv_[ib + pidx] = r.offs_c + d;
        }
      srcs_.push_back(uhmk_source{cf, ib});
    }
    if ((int)srcs_.size() - src_beg_[f] > 16)
      throw std::runtime_error("build_symbolic: more than 15 children per element is not supported");
  }
  src_beg_[nf] = (int)srcs_.size();

  // leaf copy-in maps in the element's own node order (ascending creation
  // index), which is the order the synthetic generators use for K_e.
  lmap_.clear();
  lmap_off_.assign(nf, -1);
  gen_items_.clear();
  for (int f = 0; f < nf; ++f) {
    Element_* e = order_[f];
    if (!e->is_leaf()) continue;
    lmap_off_[f] = (int)lmap_.size();
    int m = 0;
    // local position of every node of the element in the front
    for (auto& nf2 : e->nodes) {
      Node_* nd = nf2.first;
      if (!nd->n_dof) continue;
      int base = -1;
      for (auto& p : e->factor)
        if (p.first == nd) base = p.second;
      if (base < 0)
        for (auto& p : e->schur)
          if (p.first == nd) base = e->fs + p.second;
      if (base < 0) throw std::runtime_error("build_symbolic: leaf node without front position");
      for (int d = 0; d < nd->n_dof; ++d) lmap_.push_back(base + d);
      m += nd->n_dof;
    }
    if (m != fr_[f].n) throw std::runtime_error("build_symbolic: leaf dof count mismatch");
    gen_items_.push_back(uhmk_item{f, e->id, lmap_off_[f], m});
  }
  cin_blocks_.clear();
  cin_vals_.clear();
  cin_maps_.clear();
  cin_touched_.assign(2 * (size_t)nf, 0);
}

size_t FrontSolver::device_bytes() const {
  return d_arena_.bytes + d_arena32_.bytes + d_rf_.bytes + d_stage_.bytes + d_x_.bytes + d_b_.bytes + d_piv_.bytes + d_items_.bytes +
         d_val_.bytes + d_colidx_.bytes + d_cidx_.bytes + d_inv_.bytes + d_emap_.bytes;
}

void FrontSolver::create_buffers() {
  if (created_) return;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    throw std::runtime_error("no HIP device: the multifrontal path runs on the GPU only");
  check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  check_hip(hipStreamCreateWithFlags(&side_stream_, hipStreamNonBlocking), "hipStreamCreate");
  dF_.upload(fr_, stream_);
  d_emap_.upload(emap_, stream_);
  d_inv_.upload(inv_, stream_);
  d_srcbeg_.upload(src_beg_, stream_);
  d_srcs_.upload(srcs_, stream_);
  d_lmap_.upload(lmap_, stream_);
  d_goff_.upload(goff_, stream_);
  d_gen_items_.upload(gen_items_, stream_);
  std::vector<int> lv;
  level_lists_off_.assign(levels_.size(), {});
  for (auto& l : levels_) lv.insert(lv.end(), l.begin(), l.end());
  d_lvl_.upload(lv, stream_);
  d_arena_.alloc(std::max(1LL, arena_elems_) * sizeof(double));
  d_arena32_.alloc(std::max(1LL, arena32_elems_) * sizeof(float));
  d_rf_.alloc(std::max(1LL, x_elems_) * sizeof(double));
  d_z_.alloc(std::max(1LL, x_elems_) * sizeof(double));
  d_serr_.alloc(sizeof(int));
  check_hip(hipMemset(d_serr_.p, 0, sizeof(int)), "solve error word");
  check_hip(hipMemsetAsync(d_rf_.p, 0, d_rf_.bytes, stream_), "memset rf");
  d_piv_.alloc(std::max(1, piv_elems_) * sizeof(int));
  d_perm_.alloc(std::max(1, piv_elems_) * sizeof(int));
  // two panel permutations per front: with the LU lookahead the next panel
  // (odd / even by panel index) runs while the row kernel reads this one's
  d_pperm_.alloc(std::max<size_t>(1, fr_.size()) * 2 * 257 * sizeof(int));
  {
    std::vector<int> fsl;
    for (int f = 0; f < (int)fr_.size(); ++f)
      if (fr_[f].fs > 0) fsl.push_back(f);
    n_fsl_ = (int)fsl.size();
    d_fsl_.upload(fsl, stream_);
  }
  d_x_.alloc(x_elems_ * sizeof(double));
  d_b_.alloc(x_elems_ * sizeof(double));
  d_stage_.alloc(std::max(1LL, stage_elems_) * sizeof(double));
  d_info_.alloc(std::max<size_t>(1, fr_.size()) * sizeof(int));
  check_hip(hipMemsetAsync(d_stage_.p, 0, d_stage_.bytes, stream_), "memset stage");
  check_hip(hipMemsetAsync(d_b_.p, 0, d_b_.bytes, stream_), "memset b");
  check_hip(hipMemsetAsync(d_x_.p, 0, d_x_.bytes, stream_), "memset x");
  check_hip(hipMemsetAsync(d_piv_.p, 0, d_piv_.bytes, stream_), "memset piv");
  check_hip(hipStreamSynchronize(stream_), "create_buffers");
  created_ = true;
}

// ------------------------------------------------------------ staging
void FrontSolver::stage_lhs(int f, int m, const int* map, const double* K) {
  // interf.cxx:108-204 -> Helper_::merge_A into a zeroed leaf front:
  // F[map[i], map[j]] += K[i, j], in call order, on the device at the next
  // upload.
  Element_* e = order_[f];
  if (!e->is_leaf()) throw std::runtime_error("copy_in: element matrices go to leaf elements only");
  const int n = fr_[f].n;
  for (int i = 0; i < m; ++i)
    if (map[i] < 0 || map[i] >= n) throw std::runtime_error("copy_in: map outside the front");
  cin_blocks_.push_back(uhmk_cblock{(long long)cin_vals_.size(), (int)cin_maps_.size(), m, m, f, 0});
  cin_vals_.insert(cin_vals_.end(), K, K + (size_t)m * m);
  cin_maps_.insert(cin_maps_.end(), map, map + m);
  stage_dirty_ = true;
  vals_dirty_ = true;
}

void FrontSolver::stage_rhs(int f, int m, int nr, const int* map, const double* b) {
  // Helper_::merge_rhs_b then set_rhs(true) on the element (interf.cxx:193).
  Element_* e = order_[f];
  if (!e->is_leaf()) throw std::runtime_error("copy_in: right-hand sides go to leaf elements only");
  if (nr != nrhs_) throw std::runtime_error("copy_in: rhs column count differs from n_rhs");
  const int n = fr_[f].n;
  for (int i = 0; i < m; ++i)
    if (map[i] < 0 || map[i] >= n) throw std::runtime_error("copy_in: map outside the front");
  cin_blocks_.push_back(uhmk_cblock{(long long)cin_vals_.size(), (int)cin_maps_.size(), m, nr, f, 1});
  cin_vals_.insert(cin_vals_.end(), b, b + (size_t)m * nr);
  cin_maps_.insert(cin_maps_.end(), map, map + m);
  rhs_dirty_ = true;
  bg_dirty_ = true;
}

void FrontSolver::flush_copy_in() {
  if (cin_blocks_.empty()) {
    cin_vals_.clear();
    cin_maps_.clear();
    stage_dirty_ = rhs_dirty_ = false;
    return;
  }
  // one workgroup per (front, side), its blocks in call order; a front's
  // first block since creation / the last generate zeroes it first
  std::vector<int> ord(cin_blocks_.size());
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
    const uhmk_cblock &x = cin_blocks_[a], &y = cin_blocks_[b];
    return x.side != y.side ? x.side < y.side : x.f < y.f;
  });
  std::vector<uhmk_cblock> blocks;
  std::vector<uhmk_item> items;
  for (size_t q = 0; q < ord.size(); ++q) {
    const uhmk_cblock& c = cin_blocks_[ord[q]];
    if (items.empty() || items.back().f != c.f || (items.back().c & 1) != c.side) {
      char& t = cin_touched_[2 * (size_t)c.f + c.side];
      items.push_back(uhmk_item{c.f, (int)blocks.size(), 0, c.side | (t ? 0 : 2)});
      t = 1;
    }
    ++items.back().b;
    blocks.push_back(c);
  }
  d_cin_vals_.upload(cin_vals_, stream_);
  d_cin_maps_.upload(cin_maps_, stream_);
  d_cin_blocks_.upload(blocks, stream_);
  d_cin_items_.upload(items, stream_);
  check_k(uhmk_copy_in(dF_.as<uhmk_front>(), d_cin_items_.as<uhmk_item>(), (int)items.size(),
                       d_cin_blocks_.as<uhmk_cblock>(), d_cin_vals_.as<double>(), d_cin_maps_.as<int>(),
                       d_stage_.as<double>(), d_b_.as<double>(), nrhs_, stream_),
          "copy in");
  check_hip(hipStreamSynchronize(stream_), "copy in");
  cin_blocks_.clear();
  cin_vals_.clear();
  cin_maps_.clear();
  stage_dirty_ = rhs_dirty_ = false;
}

void FrontSolver::upload_stage() {
  if (stage_dirty_) flush_copy_in();
}

void FrontSolver::upload_rhs() {
  if (rhs_dirty_) flush_copy_in();
}

void FrontSolver::set_rhs() {
  // Matrix_FLA_::set_rhs (fla/matrix.cxx:325-356): x <- b at leaves, x and b
  // zero elsewhere.  b of non-leaves is never written here; x is produced
  // by the forward solve from b and the children.
  if (!created_) create_buffers();
  upload_rhs();
}

void FrontSolver::generate(unsigned long long seed, double alpha, double beta, int scale_kind, double p0,
                           double p1, int with_lhs, int with_rhs) {
  if (!created_) create_buffers();
  int max_m = 0;
  for (auto& it : gen_items_) max_m = std::max(max_m, it.c);
  check_k(uhmk_generate_elements(dF_.as<uhmk_front>(), d_gen_items_.as<uhmk_item>(), (int)gen_items_.size(), max_m,
                                 d_lmap_.as<int>(), seed, alpha, beta, scale_kind, p0, p1, d_stage_.as<double>(),
                                 d_b_.as<double>(), nrhs_, with_lhs, with_rhs, stream_),
          "generate");
  check_hip(hipStreamSynchronize(stream_), "generate");
  // the generated values replace the stages / right-hand sides: copy-in
  // blocks not yet scattered are dropped, later ones start from zero
  auto drop_side = [&](int side) {
    std::vector<uhmk_cblock> keep;
    for (const uhmk_cblock& c : cin_blocks_)
      if (c.side != side) keep.push_back(c);
    cin_blocks_.swap(keep);
    for (size_t f = 0; f < fr_.size(); ++f) cin_touched_[2 * f + side] = 0;
  };
  if (with_lhs) {
    drop_side(0);
    stage_dirty_ = false;
    vals_dirty_ = true;
  }
  if (with_rhs) {
    drop_side(1);
    rhs_dirty_ = false;
    bg_dirty_ = true;
  }
}

// ------------------------------------------------------------ schedule
static void tile_starts(int a, int b, int cut, int T, std::vector<int>& out) {
  out.clear();
  for (int s = a; s < std::min(b, cut); s += T) out.push_back(s);
  for (int s = std::max(a, cut); s < b; s += T) out.push_back(s);
}

// GEMM tile edge of one launch from the mean extent of the updated blocks
// (rows x columns): 128 x 128 tiles when both average at least the
// threshold (256 by default, tunable), so the ragged edge tiles stay a small
// share, 64 x 64 otherwise.
static int gemm_tile_for(double mean_rows, double mean_cols, double threshold) {
  return (mean_rows >= threshold && mean_cols >= threshold) ? 128 : 64;
}

void FrontSolver::build_schedule(int method) {
  const bool chol = (method == UHM_CHOL);
  items_.clear();
  flists_.clear();
  launches_.clear();
  flag_slots_ = 1;
  const int nb = nb_;
  std::vector<int> R, C;
  auto tile_end = [](int s, int fs, int n, int T) { return std::min(s + T, s < fs ? fs : n); };
  // Cholesky lookahead: the update of the next panel's column block is
  // split off each panel's GEMM launch; the next panel's diagonal block and
  // TRSM then run on the lookahead stream while the main stream does the
  // rest of the update.  Events order the two streams; every main-stream
  // launch after a lookahead pair waits for it (pending).
  int n_events = 0, pending = -1, ev_next = -1;
  bool split_prev = false, fused_prev = false;
  auto push_main = [&](Launch l) {
    l.stream = 0;
    if (pending >= 0) {
      l.wait_ev = pending;
      pending = -1;
    }
    launches_.push_back(l);
  };
  // lower-triangle tiles of front f: rows [r0, n), columns [cb, ce), edge T
  auto lower_tiles = [&](int f, int r0, int cb, int ce, int T) {
    const int fs = fr_[f].fs, n = fr_[f].n;
    tile_starts(r0, n, fs, T, R);
    tile_starts(cb, ce, fs, T, C);
    for (int j0 : C)
      for (int i0 : R) {
        if (tile_end(i0, fs, n, T) - 1 < j0) continue;
        items_.push_back(uhmk_item{f, i0, j0, 0});
      }
  };
  // the fronts to factor: all but those kept from a previous pattern
  flv_.assign(levels_.size(), {});
  for (size_t L = 0; L < levels_.size(); ++L)
    for (int f : levels_[L])
      if (reuse_.empty() || !reuse_[f]) flv_[L].push_back(f);
  std::vector<char> lp_front(fr_.size(), 0);  // fronts with low-precision Schur products
  for (int L = 0; L < (int)levels_.size(); ++L) {
    const auto& lv = flv_[L];
    if (chol && nb == 64 && small_fused_ > 0 && (int)lv.size() >= small_fused_) {
      // a level of many small fronts (every fs <= 64): one launch runs the
      // assembly, diagonal block, TRSM and Schur complement of each front
      bool small = true;
      for (int f : lv) small = small && fr_[f].fs <= 64;
      if (small) {
        Launch lf;
        lf.kind = Launch::CHOL_SMALL;
        lf.level = L;
        lf.off = (long long)flists_.size();
        lf.count = (int)lv.size();
        flists_.insert(flists_.end(), lv.begin(), lv.end());
        push_main(lf);
        continue;
      }
    }
    Launch la;
    la.kind = Launch::ASSEMBLE;
    la.level = L;
    la.off = (long long)items_.size();
    for (int f : lv) {
      const int n = fr_[f].n;
      for (int j0 = 0; j0 < n; j0 += 64)
        for (int i0 = 0; i0 < n; i0 += 64) {
          if (chol && std::min(i0 + 64, n) - 1 < j0) continue;
          items_.push_back(uhmk_item{f, i0, j0, 0});
        }
    }
    la.count = (int)(items_.size() - la.off);
    if (la.count) push_main(la);
    // LU fronts that fit one wavefront factor in one whole-front launch;
    // the panel / row / GEMM launches below cover the others
    std::vector<int> all_big;
    {
      std::vector<int> small;
      for (int f : lv) {
        if (!chol && fr_[f].fs > 0 && fr_[f].n <= std::min(64, front_small_)) small.push_back(f);
        else all_big.push_back(f);
      }
      if (!small.empty()) {
        Launch lf;
        lf.kind = Launch::LU_FRONT;
        lf.level = L;
        lf.off = (long long)flists_.size();
        lf.count = (int)small.size();
        flists_.insert(flists_.end(), small.begin(), small.end());
        push_main(lf);
      }
    }
    // size classes of the level's fronts (front_class: fs <= 64, <= 512,
    // larger), each scheduled as its own set of launches with the knobs of
    // its class (panel group, GEMM tile threshold, group launches), one
    // class after the other
    std::vector<int> bucket[N_FRONT_CLASSES];
    for (int f : all_big) bucket[front_class(fr_[f].fs)].push_back(f);
    for (int cls = 0; cls < N_FRONT_CLASSES; ++cls) {
    const std::vector<int>& big = bucket[cls];
    if (big.empty()) continue;
    const double tile_threshold = cls_tile_[cls] > 0 ? cls_tile_[cls] : tile_threshold_;
    const int group_kernel = cls_gkern_[cls] >= 0 ? cls_gkern_[cls] : group_kernel_;
    int maxfs = 0;
    for (int f : big) maxfs = std::max(maxfs, fr_[f].fs);
    // Cholesky with 64-wide panels: groups of G panels; each panel updates
    // only the rest of its group (64-wide tiles that stop at the group end),
    // and the trailing matrix receives one update with K = 64 G per group,
    // so the trailing entries are read and written once per group instead
    // of once per panel.
    const int G = (chol && nb == 64) ? std::max(1, class_group(cls)) : 1;
    // trailing update of the panel group [g0, g0 + GW): rows [ge, n),
    // columns [ge, fs), K = [g0, ge), lower tiles
    auto emit_trailing = [&](int g0, int GW) {
      Launch lt;
      lt.c0 = g0;
      lt.nb = GW;
      lt.level = L;
      lt.kind = Launch::GEMM;
      lt.extra = 1;
      double sr = 0, sc = 0;
      int cnt = 0;
      for (int f : big) {
        const int ge = std::min(g0 + GW, fr_[f].fs);
        if (fr_[f].fs <= g0 || ge >= fr_[f].fs) continue;
        sr += fr_[f].n - ge;
        sc += fr_[f].fs - ge;
        ++cnt;
      }
      lt.tile = cnt ? gemm_tile_for(sr / cnt, sc / cnt, tile_threshold) : 64;
      const int T = lt.tile;
      lt.off = (long long)items_.size();
      for (int f : big) {
        const int fs = fr_[f].fs, n = fr_[f].n;
        if (fs <= g0) continue;
        const int ge = std::min(g0 + GW, fs);
        if (ge >= fs) continue;
        tile_starts(ge, n, fs, T, R);
        tile_starts(ge, fs, fs, T, C);
        for (int j0 : C)
          for (int i0 : R) {
            if (tile_end(i0, fs, n, T) - 1 < j0) continue;
            items_.push_back(uhmk_item{f, i0, j0, 0});
          }
      }
      lt.count = (int)(items_.size() - lt.off);
      if (lt.count) push_main(lt);
    };
    // group launches where they measured faster than per-panel launches:
    // levels with few fronts (the top of the tree, whose panel chains are
    // long); below, the many small fronts of a level fill the chip better
    // with one launch per panel step
    const bool group_level = group_kernel == 1 || (group_kernel == 2 && (int)big.size() <= 16);
    if (chol && nb == 64 && group_level) {
      // each group of G panels in one persistent launch (uhmk_chol_group):
      // per front one item per 64-row tile of rows [g0, n), the tiles of
      // the group's diagonal range of every front first, then the trailing
      // update of the group
      const int GW = 64 * std::min(G, 16);
      // workgroups resident at once (two per CU at the launch's LDS size):
      // a launch that would exceed them gives each workgroup below a front's
      // diagonal range several consecutive row tiles, since those workgroups
      // hold their slot until the group's last panel
      int cus = 256;
      {
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess)
          (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      }
      const int slots = 2 * std::max(1, cus);
      for (int g0 = 0; g0 < maxfs; g0 += GW) {
        Launch lk;
        lk.kind = Launch::CHOL_GROUP;
        lk.c0 = g0;
        lk.nb = GW;
        lk.level = L;
        lk.off = (long long)items_.size();
        int n_diag = 0, n_below = 0;
        for (int f : big) {
          const int fs = fr_[f].fs, n = fr_[f].n;
          if (fs <= g0) continue;
          const int P = (std::min(g0 + GW, fs) - g0 + 63) / 64;
          const int rows = (fs - g0 + 63) / 64 + (n - fs + 63) / 64;
          n_diag += P;
          n_below += rows - P;
        }
        const int per = (n_diag < slots && n_diag + n_below > slots)
                            ? std::min(16, (n_below + (slots - n_diag) - 1) / (slots - n_diag))
                            : 1;
        std::vector<uhmk_item> below;
        for (int f : big) {
          const int fs = fr_[f].fs, n = fr_[f].n;
          if (fs <= g0) continue;
          const int P = (std::min(g0 + GW, fs) - g0 + 63) / 64;
          const int rows = (fs - g0 + 63) / 64 + (n - fs + 63) / 64;
          const int slot = flag_slots_;
          flag_slots_ += 16;
          for (int r = 0; r < P; ++r) items_.push_back(uhmk_item{f, r, 1, slot});
          for (int r = P; r < rows; r += per) below.push_back(uhmk_item{f, r, std::min(per, rows - r), slot});
        }
        items_.insert(items_.end(), below.begin(), below.end());
        lk.count = (int)(items_.size() - lk.off);
        if (lk.count) push_main(lk);
        emit_trailing(g0, GW);
      }
      maxfs = 0;  // the panel loop below has nothing left to do
    }
    for (int c0 = 0; c0 < maxfs; c0 += nb) {
      const int g0 = c0 - c0 % (G * nb);  // first column of this panel's group
      const bool group_last = (c0 + nb >= g0 + G * nb) || (c0 + nb >= maxfs);
      std::vector<int> act;
      int maxrows = 0;
      for (int f : big)
        if (fr_[f].fs > c0) {
          act.push_back(f);
          maxrows = std::max(maxrows, fr_[f].fs - c0);
        }
      Launch lp;
      lp.c0 = c0;
      lp.nb = nb;
      lp.level = L;
      lp.kind = chol ? Launch::CHOL_DIAG : Launch::LU_PANEL;
      // the previous panel's update factored every diagonal block of this one
      const bool diag_done = fused_prev;
      fused_prev = false;
      lp.off = (long long)flists_.size();
      lp.count = (int)act.size();
      lp.extra = maxrows;
      flists_.insert(flists_.end(), act.begin(), act.end());

      Launch lr;
      lr.c0 = c0;
      lr.nb = nb;
      lr.level = L;
      lr.kind = chol ? Launch::CHOL_TRSM : Launch::LU_ROWS;
      lr.off = (long long)items_.size();
      const int rw = chol ? 64 : lu_rows_w_;  // LU row-kernel chunk (tile field of the launch)
      lr.tile = rw;
      for (int f : act) {
        const int fs = fr_[f].fs, n = fr_[f].n;
        const int c1 = std::min(c0 + nb, fs);
        if (chol) {
          for (int i0 = c1; i0 < n; i0 += 64) items_.push_back(uhmk_item{f, 0, i0, 0});
        } else {
          for (int j0 = 0; j0 < c0; j0 += rw) items_.push_back(uhmk_item{f, 0, j0, 0});
          for (int i0 = fs; i0 < n; i0 += rw) items_.push_back(uhmk_item{f, 1, i0, 0});
          for (int j0 = c1; j0 < n; j0 += rw) items_.push_back(uhmk_item{f, 2, j0, 0});
        }
      }
      lr.count = (int)(items_.size() - lr.off);
      // LU lookahead: the next panel's column block (row kernel and ATL
      // update) goes first; the next panel then runs on the lookahead stream
      // while the main stream applies this panel to the rest
      const bool lla = !chol && lu_lookahead_ && nb == 64 && c0 + nb < maxfs;
      if (split_prev && !chol) {
        // this LU panel on the lookahead stream, after the update of its
        // columns (ev_next); the main stream waits for it before its row kernel
        lp.stream = 1;
        lp.wait_ev = ev_next;
        lp.record_ev = pending = n_events++;
        launches_.push_back(lp);
        if (lr.count && !lla) push_main(lr);
      } else if (split_prev) {
        // this panel's diagonal block and TRSM on the lookahead stream, after
        // the update of its columns (ev_next); the main stream waits for them
        // before its next launch
        lp.stream = 1;
        lp.wait_ev = ev_next;
        launches_.push_back(lp);
        if (lr.count) {
          lr.stream = 1;
          launches_.push_back(lr);
        }
        launches_.back().record_ev = pending = n_events++;
      } else {
        if (!diag_done) push_main(lp);
        if (lr.count && !lla) push_main(lr);
      }
      split_prev = false;
      const bool la = chol && lookahead_ && c0 + nb < maxfs;
      Launch lr_rest = lr;
      if (lla) {
        // row-kernel items of the next panel's block (kind 2 at c1) first
        auto next_block = [&](const uhmk_item& it) {
          const int c1 = std::min(c0 + nb, fr_[it.f].fs);
          return it.a == 2 && it.b >= c1 && it.b < c1 + 64;
        };
        std::stable_partition(items_.begin() + lr.off, items_.end(), next_block);
        int nn = 0;
        for (long long q = lr.off; q < (long long)items_.size() && next_block(items_[q]); ++q) ++nn;
        lr.count = nn;
        lr_rest.off = lr.off + nn;
        lr_rest.count -= nn;
        if (lr.count) push_main(lr);
      }

      if (G > 1) {
        // update of the rest of the group by this panel (K = [c0, c0+nb))
        Launch li;
        li.c0 = c0;
        li.nb = nb;
        li.level = L;
        li.kind = Launch::GEMM;
        li.extra = 1;
        li.tile = 64;
        li.off = (long long)items_.size();
        for (int f : act) {
          const int fs = fr_[f].fs, n = fr_[f].n;
          const int c1 = std::min(c0 + nb, fs), ge = std::min(g0 + G * nb, fs);
          if (c1 >= ge) continue;
          tile_starts(c1, n, fs, 64, R);
          tile_starts(c1, ge, fs, 64, C);
          for (int j0 : C)
            for (int i0 : R) {
              if (tile_end(i0, fs, n, 64) - 1 < j0) continue;
              items_.push_back(uhmk_item{f, i0, j0, 0});
            }
        }
        li.count = (int)(items_.size() - li.off);
        if (la && !group_last && li.count) {
          // the next panel's column block (first 64-wide column tile) first
          std::stable_partition(items_.begin() + li.off, items_.end(), [&](const uhmk_item& it) {
            return it.b == std::min(c0 + nb, fr_[it.f].fs);
          });
          int nn = 0;
          for (long long q = li.off; q < (long long)items_.size(); ++q)
            if (items_[q].b == std::min(c0 + nb, fr_[items_[q].f].fs)) ++nn;
          Launch rest = li;
          li.count = nn;
          rest.off = li.off + nn;
          rest.count -= nn;
          if (li.count) {
            li.record_ev = ev_next = n_events++;
            push_main(li);
            split_prev = true;
          }
          if (rest.count) push_main(rest);
        } else if (li.count) {
          if (fuse_diag_ && !la && !group_last && nb == 64) {
            // the next panel's diagonal tiles (c1, c1) first, flagged: their
            // workgroups factor the block after updating it
            auto diag_tile = [&](const uhmk_item& it) {
              const int c1 = std::min(c0 + nb, fr_[it.f].fs);
              return it.a == c1 && it.b == c1;
            };
            std::stable_partition(items_.begin() + li.off, items_.end(), diag_tile);
            for (long long q = li.off; q < (long long)items_.size() && diag_tile(items_[q]); ++q) items_[q].c = 1;
            li.kind = Launch::GEMM_DIAG;
            fused_prev = true;
          }
          push_main(li);
        }
        if (!group_last) continue;
        // trailing update of the whole group (K = [g0, g0 + G nb))
        Launch lt;
        lt.c0 = g0;
        lt.nb = G * nb;
        lt.level = L;
        lt.kind = Launch::GEMM;
        lt.extra = 1;
        {
          double sr = 0, sc = 0;
          int cnt = 0;
          for (int f : big) {
            const int ge = std::min(g0 + G * nb, fr_[f].fs);
            if (fr_[f].fs <= g0 || ge >= fr_[f].fs) continue;
            sr += fr_[f].n - ge;
            sc += fr_[f].fs - ge;
            ++cnt;
          }
          lt.tile = cnt ? gemm_tile_for(sr / cnt, sc / cnt, tile_threshold) : 64;
        }
        const int T = lt.tile;
        if (la) {
          // the next panel's column block [ge, ge + 64) in 64-wide tiles
          // first, then the rest of the trailing matrix
          Launch ln = lt;
          ln.tile = 64;
          ln.off = (long long)items_.size();
          for (int f : big) {
            const int fs = fr_[f].fs;
            if (fs <= g0) continue;
            const int ge = std::min(g0 + G * nb, fs);
            if (ge >= fs) continue;
            lower_tiles(f, ge, ge, std::min(ge + 64, fs), 64);
          }
          ln.count = (int)(items_.size() - ln.off);
          if (ln.count) {
            ln.record_ev = ev_next = n_events++;
            push_main(ln);
            split_prev = true;
          }
          lt.off = (long long)items_.size();
          for (int f : big) {
            const int fs = fr_[f].fs;
            if (fs <= g0) continue;
            const int ge = std::min(g0 + G * nb, fs);
            if (ge + 64 >= fs) continue;
            lower_tiles(f, ge + 64, ge + 64, fs, T);
          }
          lt.count = (int)(items_.size() - lt.off);
          if (lt.count) push_main(lt);
          continue;
        }
        lt.off = (long long)items_.size();
        for (int f : big) {
          const int fs = fr_[f].fs, n = fr_[f].n;
          if (fs <= g0) continue;
          const int ge = std::min(g0 + G * nb, fs);
          if (ge >= fs) continue;
          tile_starts(ge, n, fs, T, R);
          tile_starts(ge, fs, fs, T, C);
          for (int j0 : C)
            for (int i0 : R) {
              if (tile_end(i0, fs, n, T) - 1 < j0) continue;
              items_.push_back(uhmk_item{f, i0, j0, 0});
            }
        }
        lt.count = (int)(items_.size() - lt.off);
        if (lt.count) push_main(lt);
        continue;
      }
      Launch lg;
      lg.c0 = c0;
      lg.nb = nb;
      lg.level = L;
      lg.kind = Launch::GEMM;
      lg.extra = chol ? 1 : 0;
      {
        double sr = 0, sc = 0;
        for (int f : act) {
          const int c1 = std::min(c0 + nb, fr_[f].fs);
          sr += fr_[f].n - c1;
          sc += (chol ? fr_[f].fs : fr_[f].n) - c1;
        }
        lg.tile = gemm_tile_for(sr / act.size(), sc / act.size(), tile_threshold);
      }
      const int T = lg.tile;
      if (la) {
        // (Cholesky) the next panel's column block in 64-wide tiles first
        Launch ln = lg;
        ln.tile = 64;
        ln.off = (long long)items_.size();
        for (int f : act) {
          const int fs = fr_[f].fs, c1 = std::min(c0 + nb, fs);
          if (c1 >= fs) continue;
          lower_tiles(f, c1, c1, std::min(c1 + 64, fs), 64);
        }
        ln.count = (int)(items_.size() - ln.off);
        if (ln.count) {
          ln.record_ev = ev_next = n_events++;
          push_main(ln);
          split_prev = true;
        }
        lg.off = (long long)items_.size();
        for (int f : act) {
          const int fs = fr_[f].fs, c1 = std::min(c0 + nb, fs);
          if (c1 + 64 >= fs) continue;
          lower_tiles(f, c1 + 64, c1 + 64, fs, T);
        }
        lg.count = (int)(items_.size() - lg.off);
        if (lg.count) push_main(lg);
        continue;
      }
      if (lla) {
        // (LU) the ATL rows of the next panel's column block in 64 x 64
        // tiles, then the row kernel's other items and the rest of the update
        Launch ln = lg;
        ln.tile = 64;
        ln.off = (long long)items_.size();
        for (int f : act) {
          const int fs = fr_[f].fs, c1 = std::min(c0 + nb, fs);
          for (int i0 = c1; i0 < fs; i0 += 64) items_.push_back(uhmk_item{f, i0, c1, 0});
        }
        ln.count = (int)(items_.size() - ln.off);
        ln.record_ev = ev_next = n_events++;
        push_main(ln);
        split_prev = true;
        if (lr_rest.count) push_main(lr_rest);
        lg.tile = 64;
        lg.off = (long long)items_.size();
        for (int f : act) {
          const int fs = fr_[f].fs, n = fr_[f].n;
          const int c1 = std::min(c0 + nb, fs);
          tile_starts(c1, n, fs, 64, R);
          tile_starts(c1, n, fs, 64, C);
          for (int j0 : C)
            for (int i0 : R) {
              if (i0 >= fs && j0 >= fs) continue;
              if (i0 < fs && j0 == c1) continue;  // in the lookahead launch
              items_.push_back(uhmk_item{f, i0, j0, 0});
            }
        }
        lg.count = (int)(items_.size() - lg.off);
        if (lg.count) push_main(lg);
        continue;
      }
      lg.off = (long long)items_.size();
      for (int f : act) {
        const int fs = fr_[f].fs, n = fr_[f].n;
        const int c1 = std::min(c0 + nb, fs);
        tile_starts(c1, n, fs, T, R);
        tile_starts(c1, chol ? fs : n, fs, T, C);
        for (int j0 : C)
          for (int i0 : R) {
            if (!chol && i0 >= fs && j0 >= fs) continue;
            if (chol && tile_end(i0, fs, n, T) - 1 < j0) continue;
            items_.push_back(uhmk_item{f, i0, j0, 0});
          }
      }
      lg.count = (int)(items_.size() - lg.off);
      if (lg.count) push_main(lg);
    }
    Launch ls;
    ls.kind = Launch::SCHUR;
    ls.level = L;
    ls.extra = chol ? 1 : 0;
    bool lowp = prec_mode_ >= 2;
    {
      double sss = 0;
      int cnt = 0;
      for (int f : big)
        if (fr_[f].fs && fr_[f].n > fr_[f].fs) {
          sss += fr_[f].n - fr_[f].fs;
          ++cnt;
        }
      // low-precision operands only where the Schur blocks are large: on
      // small ones the packing saves next to nothing and the rounding costs
      // refinement steps
      if (lowp && (!cnt || sss / cnt < lp_min_ss_)) lowp = false;
      ls.mma = lowp ? prec_mode_ : -1;
      if (lowp) ls.kind = Launch::SCHUR_LP;
      if (lowp)
        for (int f : big) lp_front[f] = 1;
      // the packed low-precision products run 64-wide tiles at every size
      // (eight workgroups per CU hide their load latency)
      ls.tile = cnt && !lowp ? gemm_tile_for(sss / cnt, sss / cnt, tile_threshold) : 64;
    }
    if (lowp) {
      // pack the operands of every front of the level (64 x 64 tiles over
      // the padded packed rows and K)
      Launch lp;
      lp.kind = Launch::LP_PACK;
      lp.level = L;
      lp.off = (long long)items_.size();
      for (int f : big) {
        const int fs = fr_[f].fs, ss = fr_[f].n - fs;
        if (!fs || !ss) continue;
        const int rt = ((ss + 127) / 128) * 2, kt = (fs + 63) / 64;
        for (int op = 0; op < (chol ? 1 : 2); ++op)
          for (int a = 0; a < rt; ++a)
            for (int b = 0; b < kt; ++b) items_.push_back(uhmk_item{f, a, b, op});
      }
      lp.count = (int)(items_.size() - lp.off);
      if (lp.count) push_main(lp);
    }
    const int T = ls.tile;
    ls.off = (long long)items_.size();
    for (int f : big) {
      const int fs = fr_[f].fs, n = fr_[f].n;
      if (!fs || n == fs) continue;
      for (int j0 = fs; j0 < n; j0 += T)
        for (int i0 = fs; i0 < n; i0 += T) {
          if (chol && std::min(i0 + T, n) - 1 < j0) continue;
          items_.push_back(uhmk_item{f, i0, j0, 0});
        }
    }
    ls.count = (int)(items_.size() - ls.off);
    if (ls.count) push_main(ls);
    }  // size classes
  }
  final_wait_ev_ = pending;
  while ((int)la_ev_.size() < n_events) {
    hipEvent_t e;
    check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
    la_ev_.push_back(e);
  }
  for (Launch& l : launches_) {
    l.prec = level_prec_[l.level];
    // mode 2: the Schur complement products take BF16 operands; the panel
    // (trailing) updates, which form the factor columns themselves, stay FP32
    // mma 2 / 3: packed BF16 / FP16 Schur products
    l.mma = prec_mode_ == 0 ? 0 : (l.kind == Launch::SCHUR_LP ? prec_mode_ : 1);
    // the group launches' small in-group products in the storage type (FP64
    // products on FP64 fronts measured faster there than FP32 ones)
    if (l.kind == Launch::CHOL_GROUP) l.mma = l.prec == 0 ? 0 : 1;
    // whole small fronts: the Schur products of the mode (FP32 in the
    // low-precision modes, whose packing does not pay on fronts this small)
    if (l.kind == Launch::CHOL_SMALL) l.mma = prec_mode_ == 0 ? 0 : 1;
  }
  // Cholesky, FP32 products on FP64 fronts: every Schur and trailing-update
  // launch of those levels reads an FP32 copy of its operand (element loads
  // of FP64 entries were the slow part of these products), made by a pack
  // launch just before it.  One launch's copies at a time: the launches run
  // in order on the main stream and share one buffer.
  long long p32_elems = 0;
  if (chol && schur_p32_ && prec_mode_ >= 1) {
    std::vector<Launch> out;
    out.reserve(launches_.size() * 2);
    for (const Launch& l0 : launches_) {
      Launch l = l0;
      const bool schur = l.kind == Launch::SCHUR;
      const bool conv = (schur || l.kind == Launch::GEMM) && l.extra == 1 && l.prec == 0 && l.mma == 1 &&
                        l.tile == 64 && l.stream == 0 && l.wait_ev < 0 && l.record_ev < 0 && l.count > 0;
      if (!conv) {
        out.push_back(l);
        continue;
      }
      std::map<int, long long> off;  // front -> copy offset / 64
      long long tot = 0;
      for (long long q = l.off; q < l.off + l.count; ++q) {
        const int f = items_[q].f;
        if (off.count(f)) continue;
        off[f] = tot / 64;
        tot += (uhmk_p32_size(fr_[f].n, fr_[f].fs, l.c0, l.nb, schur) + 63) / 64 * 64;
      }
      if (!tot || tot / 64 > std::numeric_limits<int>::max()) {
        out.push_back(l);
        continue;
      }
      p32_elems = std::max(p32_elems, tot);
      Launch pk = l;
      pk.kind = Launch::P32_PACK;
      pk.extra = schur ? 1 : 0;
      pk.off = (long long)items_.size();
      for (const auto& fo : off) {
        const int f = fo.first, fs = fr_[f].fs, n = fr_[f].n;
        const int kbeg = schur ? 0 : l.c0, kend = schur ? fs : std::min(l.c0 + l.nb, fs);
        for (int a = 0; a < (n - kend + 63) / 64; ++a)
          for (int b = 0; b < (kend - kbeg + 63) / 64; ++b) items_.push_back(uhmk_item{f, a, b, (int)fo.second});
      }
      pk.count = (int)(items_.size() - pk.off);
      out.push_back(pk);
      for (long long q = l.off; q < l.off + l.count; ++q) items_[q].c = (int)off[items_[q].f];
      l.kind = schur ? Launch::SCHUR_P32 : Launch::GEMM_P32;
      out.push_back(l);
    }
    launches_.swap(out);
  }
  // Algorithmic work of every launch: the flops / bytes the operation needs,
  // not what a kernel happens to execute (diagonal SYRK tiles count their
  // lower half, assembly counts each stored entry once).
  auto tri = [](double m) { return 0.5 * m * (m + 1.0); };
  for (Launch& l : launches_) {
    double fl = 0, by = 0;
    switch (l.kind) {
      case Launch::ASSEMBLE:
        for (int f : flv_[l.level]) {
          const double n = fr_[f].n;
          const double w = chol ? tri(n) : n * n;
          double rd = fr_[f].soff >= 0 ? w : 0.0;
          for (Element_* c : order_[f]->children) rd += chol ? tri(c->ss) : (double)c->ss * c->ss;
          by += 8.0 * (w + rd);
        }
        break;
      case Launch::LU_PANEL:
      case Launch::CHOL_DIAG:
        for (int q = 0; q < l.count; ++q) {
          const int f = flists_[l.off + q];
          const double m = fr_[f].fs - l.c0, kb = std::min<double>(l.nb, m);
          if (l.kind == Launch::LU_PANEL) {
            fl += flop_lu(m, kb);
            by += 16.0 * m * kb;
          } else {
            fl += flop_chol(kb);
            by += 16.0 * kb * kb;
          }
        }
        break;
      case Launch::LU_ROWS:
      case Launch::CHOL_TRSM:
        for (int q = 0; q < l.count; ++q) {
          const uhmk_item& it = items_[l.off + q];
          const int fs = fr_[it.f].fs, n = fr_[it.f].n;
          const double kb = std::min(l.nb, fs - l.c0);
          const int cw = l.kind == Launch::LU_ROWS ? l.tile : 64;
          if (l.kind == Launch::CHOL_TRSM || it.a == 1) {
            const double h = std::min(cw, n - it.b);
            fl += kb * kb * h;
            by += 16.0 * kb * h;
          } else {
            const double w = std::min(cw, (it.a == 0 ? l.c0 : n) - it.b);
            if (it.a == 2) fl += kb * kb * w;
            by += 16.0 * 2.0 * kb * w;
          }
        }
        break;
      case Launch::LP_PACK:
        for (int q = 0; q < l.count; ++q) {
          const uhmk_item& it = items_[l.off + q];
          const int fs = fr_[it.f].fs, ss = fr_[it.f].n - fs;
          const double rows = std::max(0, std::min(64, ss - 64 * it.a)), ks = std::max(0, std::min(64, fs - 64 * it.b));
          by += 8.0 * rows * ks + 2.0 * 64 * 64;  // read (FP64 units, halved below for FP32), write packed
        }
        break;
      case Launch::CHOL_GROUP:
        for (int q = 0; q < l.count; ++q) {
          const uhmk_item& it = items_[l.off + q];
          const int fs = fr_[it.f].fs, n = fr_[it.f].n;
          const int ge = std::min(l.c0 + l.nb, fs), P = (ge - l.c0 + 63) / 64;
          const int ratl = (fs - l.c0 + 63) / 64;
          for (int t = it.a; t < it.a + it.b; ++t) {
            const int i0 = t < ratl ? l.c0 + 64 * t : fs + 64 * (t - ratl);
            const double h = std::min(i0 + 64, i0 < fs ? fs : n) - i0;
            const int npan = t < P ? t + 1 : P;
            for (int p = 0; p < npan; ++p) {
              const int pc0 = l.c0 + 64 * p;
              const double kb = std::min(64, ge - pc0), K = pc0 - l.c0;
              if (t == p) {
                fl += kb * (kb + 1.0) * K + flop_chol(kb);
                by += 8.0 * K * kb + 16.0 * kb * kb;
              } else {
                fl += 2.0 * h * kb * K + kb * kb * h;
                by += 8.0 * K * (h + kb) + 16.0 * h * kb;
              }
            }
          }
        }
        break;
      case Launch::CHOL_SMALL:
        for (int q = 0; q < l.count; ++q) {
          const int f = flists_[l.off + q];
          const double n = fr_[f].n, fs = fr_[f].fs, ss = n - fs;
          double rd = fr_[f].soff >= 0 ? tri(n) : 0.0;
          for (Element_* c : order_[f]->children) rd += tri(c->ss);
          fl += flop_chol(fs) + fs * fs * ss + fs * ss * (ss + 1.0);
          by += 8.0 * (tri(n) + rd);
        }
        break;
      case Launch::LU_FRONT:
        for (int q = 0; q < l.count; ++q) {
          const int f = flists_[l.off + q];
          const double fs = fr_[f].fs, ss = fr_[f].n - fr_[f].fs, n = fr_[f].n;
          fl += flop_lu(fs, fs) + 2.0 * fs * fs * ss + 2.0 * fs * ss * ss;
          by += 16.0 * n * n;
        }
        break;
      case Launch::P32_PACK:
        for (int q = 0; q < l.count; ++q) {
          const uhmk_item& it = items_[l.off + q];
          const int fs = fr_[it.f].fs, n = fr_[it.f].n;
          const int kbeg = l.extra ? 0 : l.c0, kend = l.extra ? fs : std::min(l.c0 + l.nb, fs);
          const double rows = std::max(0, std::min(64, n - kend - 64 * it.a));
          const double ks = std::max(0, std::min(64, kend - kbeg - 64 * it.b));
          by += 12.0 * rows * ks;  // FP64 read, FP32 write
        }
        break;
      case Launch::GEMM:
      case Launch::GEMM_DIAG:
      case Launch::SCHUR:
      case Launch::SCHUR_LP:
      case Launch::SCHUR_P32:
      case Launch::GEMM_P32:
        for (int q = 0; q < l.count; ++q) {
          const uhmk_item& it = items_[l.off + q];
          const int fs = fr_[it.f].fs, n = fr_[it.f].n;
          if (l.kind == Launch::GEMM_DIAG && it.c == 1) fl += flop_chol(std::min(l.nb, fs - it.a));
          const double K = (l.kind == Launch::SCHUR || l.kind == Launch::SCHUR_LP || l.kind == Launch::SCHUR_P32)
                               ? fs
                               : std::min(l.c0 + l.nb, fs) - l.c0;
          const int i1 = tile_end(it.a, fs, n, l.tile), j1 = tile_end(it.b, fs, n, l.tile);
          double cnt = 0;
          if (chol)
            for (int j = it.b; j < j1; ++j) cnt += std::max(0, i1 - std::max(it.a, j));
          else
            cnt = (double)(i1 - it.a) * (j1 - it.b);
          fl += 2.0 * K * cnt;
          by += 16.0 * cnt + 8.0 * K * ((i1 - it.a) + (j1 - it.b));
        }
        break;
    }
    l.flops = fl;
    l.bytes = l.prec ? 0.5 * by : by;
  }
  d_items_.upload(items_, stream_);
  d_flists_.upload(flists_, stream_);
  d_flags_.reserve((size_t)flag_slots_ * sizeof(int));
  if (prec_mode_ >= 2) {
    // packed operands of the fronts of the levels with low-precision products
    std::vector<long long> lpoff(fr_.size(), 0);
    long long tot = 0;
    for (size_t L = 0; L < levels_.size(); ++L)
      for (int f : flv_[L]) {
        lpoff[f] = tot;
        if (lp_front[f]) tot += (uhmk_schur_lp_size(fr_[f].n, fr_[f].fs, chol ? 0 : 1) + 63) / 64 * 64;
      }
    lp_elems_ = tot;
    d_lp_.reserve(std::max(1LL, tot) * sizeof(unsigned short));
    d_lpoff_.upload(lpoff, stream_);
    d_lpamax_.reserve(std::max<size_t>(1, 2 * fr_.size()) * sizeof(unsigned));
  }
  // FP32 operand copies (offsets multiples of 64 entries: 16-byte aligned)
  if (p32_elems) d_p32_.reserve((size_t)p32_elems * sizeof(float));
  check_hip(hipStreamSynchronize(stream_), "schedule upload");
  if (factor_graph_) (void)hipGraphExecDestroy(factor_graph_);
  factor_graph_ = nullptr;
  schedule_method_ = method;
}

void FrontSolver::enqueue_factor(int method, std::vector<hipEvent_t>* pev) {
  // The level schedule as a fixed launch sequence on stream_ and, for the
  // Cholesky lookahead, side_stream_ (ordered by events; also the body that
  // a graph capture records: the side stream forks from and joins back into
  // stream_; no memset nodes, the info flags are cleared by a fill kernel).
  const uhmk_front* F = dF_.as<uhmk_front>();
  const uhmk_item* IT = d_items_.as<uhmk_item>();
  const int* FL = d_flists_.as<int>();
  int* P = d_piv_.as<int>();
  int* info = d_info_.as<int>();
  check_k(uhmk_fill_int(info, (long long)(d_info_.bytes / sizeof(int)), 0, stream_), "clear info");
  check_k(uhmk_fill_int(d_flags_.as<int>(), flag_slots_, 0, stream_), "clear group flags");
  if (prec_mode_ == 3)
    check_k(uhmk_fill_int(d_lpamax_.as<int>(), (long long)(d_lpamax_.bytes / sizeof(int)), 0, stream_),
            "clear operand maxima");
  for (size_t li = 0; li < launches_.size(); ++li) {
    const Launch& l = launches_[li];
    int rc = 0;
    hipStream_t st = l.stream ? side_stream_ : stream_;
    if (l.wait_ev >= 0) check_hip(hipStreamWaitEvent(st, la_ev_[l.wait_ev], 0), "stream wait");
    if (pev) check_hip(hipEventRecord((*pev)[2 * li], st), "event");
    void* A = arena_ptr(l.prec);
    switch (l.kind) {
      case Launch::CHOL_SMALL:
        rc = uhmk_chol_front_small(F, FL + l.off, l.count, d_srcbeg_.as<int>(), d_srcs_.as<uhmk_source>(),
                                   d_inv_.as<int>(), d_arena_.as<double>(), d_arena32_.as<float>(), l.prec, l.mma,
                                   d_stage_.as<double>(), info, st);
        break;
      case Launch::ASSEMBLE:
        rc = uhmk_assemble(F, IT + l.off, l.count, d_srcbeg_.as<int>(), d_srcs_.as<uhmk_source>(), d_inv_.as<int>(),
                           d_arena_.as<double>(), d_arena32_.as<float>(), l.prec, d_stage_.as<double>(), st);
        break;
      case Launch::LU_PANEL:
        rc = (lu_panel_ll_ ? uhmk_lu_panel_ll : uhmk_lu_panel)(F, FL + l.off, l.count, l.extra, l.c0, l.nb, A,
                                                               l.prec, P, info, pperm_of(l), st);
        break;
      case Launch::LU_ROWS:
        rc = uhmk_lu_rows(F, IT + l.off, l.count, l.c0, l.nb, A, l.prec, pperm_of(l), l.tile, st);
        break;
      case Launch::LU_FRONT:
        rc = uhmk_lu_front64(F, FL + l.off, l.count, A, l.prec, P, info, st);
        break;
      case Launch::CHOL_DIAG:
        rc = uhmk_chol_diag(F, FL + l.off, l.count, l.c0, l.nb, A, l.prec, info, st);
        break;
      case Launch::CHOL_TRSM:
        rc = uhmk_chol_trsm(F, IT + l.off, l.count, l.c0, l.nb, A, l.prec, st);
        break;
      case Launch::GEMM:
        rc = uhmk_gemm_update(F, IT + l.off, l.count, l.c0, l.nb, 0, l.extra, l.tile, A, l.prec, l.mma, st);
        break;
      case Launch::GEMM_DIAG:
        rc = uhmk_chol_gemm_diag(F, IT + l.off, l.count, l.c0, l.nb, A, l.prec, l.mma, info, st);
        break;
      case Launch::SCHUR:
      case Launch::SCHUR_LP:
        if (l.kind == Launch::SCHUR_LP)
          rc = uhmk_schur_lp(F, IT + l.off, l.count, l.extra ? 0 : 1, l.tile, l.mma == 3, A, l.prec,
                             d_lpoff_.as<long long>(), d_lpamax_.as<unsigned>(), d_lp_.as<unsigned short>(), st);
        else
          rc = uhmk_gemm_update(F, IT + l.off, l.count, 0, 0, 1, l.extra, l.tile, A, l.prec, l.mma, st);
        break;
      case Launch::LP_PACK:
        rc = uhmk_schur_lp_pack(F, IT + l.off, l.count, prec_mode_ == 3, A, l.prec, d_lpoff_.as<long long>(),
                                d_lpamax_.as<unsigned>(), d_lp_.as<unsigned short>(), st);
        break;
      case Launch::P32_PACK:
        rc = uhmk_p32_pack(F, IT + l.off, l.count, l.c0, l.nb, l.extra, A, l.prec, d_p32_.as<float>(), st);
        break;
      case Launch::SCHUR_P32:
      case Launch::GEMM_P32:
        rc = uhmk_gemm_update_p32(F, IT + l.off, l.count, l.c0, l.nb, l.kind == Launch::SCHUR_P32 ? 1 : 0, A, l.prec,
                                  d_p32_.as<float>(), st);
        break;
      case Launch::CHOL_GROUP:
        rc = uhmk_chol_group(F, IT + l.off, l.count, l.c0, l.nb, A, l.prec, l.mma, d_flags_.as<int>(), info, st);
        break;
    }
    check_k(rc, "factor");
    if (pev) check_hip(hipEventRecord((*pev)[2 * li + 1], st), "event");
    if (l.record_ev >= 0) check_hip(hipEventRecord(la_ev_[l.record_ev], st), "event");
  }
  if (final_wait_ev_ >= 0) check_hip(hipStreamWaitEvent(stream_, la_ev_[final_wait_ev_], 0), "stream wait");
  if (method == UHM_LU_PIV)
    check_k(uhmk_lu_perm(F, d_fsl_.as<int>(), n_fsl_, P, d_perm_.as<int>(), stream_), "lu perm");
}

void FrontSolver::drop_graphs() {
  if (factor_graph_) (void)hipGraphExecDestroy(factor_graph_);
  factor_graph_ = nullptr;
  for (auto& g : solve_graphs_)
    if (g.second) (void)hipGraphExecDestroy(g.second);
  solve_graphs_.clear();
}

hipGraphExec_t FrontSolver::capture(const std::function<void()>& body) {
  // Record body's launches on stream_ into a graph and instantiate it.
  hipGraph_t g = nullptr;
  check_hip(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
  try {
    body();
  } catch (...) {
    (void)hipStreamEndCapture(stream_, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  check_hip(hipStreamEndCapture(stream_, &g), "end capture");
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  check_hip(e, "graph instantiate");
  return ex;
}

void FrontSolver::set_graph(bool on) {
  if (!on) drop_graphs();
  use_graph_ = on;
}

void FrontSolver::set_tuning(int key, int value) {
  if (key == 1) {
    set_chol_group(value);
  } else if (key == 2) {
    if (value < 64) throw std::runtime_error("set_tuning: tile threshold must be >= 64");
    // (values above every front size select 64 x 64 tiles everywhere)
    if (value != tile_threshold_) schedule_method_ = 0;
    tile_threshold_ = value;
  } else if (key == 3) {
    if (value < 1) throw std::runtime_error("set_tuning: block-solve threshold must be >= 1");
    block_rhs_ = value;
  } else if (key == 4) {
    if (value < 0 || value > 64) throw std::runtime_error("set_tuning: whole-front LU limit must be 0..64");
    if (value != front_small_) schedule_method_ = 0;
    front_small_ = value;
  } else if (key == 5) {
    if (value != 0 && value != 1) throw std::runtime_error("set_tuning: lookahead is 0 or 1");
    if (value != lookahead_) schedule_method_ = 0;
    lookahead_ = value;
  } else if (key == 6) {
    if (value != 0 && value != 1) throw std::runtime_error("set_tuning: diagonal fusion is 0 or 1");
    if (value != fuse_diag_) schedule_method_ = 0;
    fuse_diag_ = value;
  } else if (key >= 100 && key < 100 + N_FRONT_CLASSES) {
    if (value < 0) throw std::runtime_error("set_tuning: class panel group takes 0 (global) or a panel count");
    if (value != cls_group_[key - 100]) schedule_method_ = 0;
    cls_group_[key - 100] = value;
  } else if (key >= 110 && key < 110 + N_FRONT_CLASSES) {
    if (value < 0) throw std::runtime_error("set_tuning: class tile threshold takes 0 (global) or an extent");
    if (value != cls_tile_[key - 110]) schedule_method_ = 0;
    cls_tile_[key - 110] = value;
  } else if (key >= 120 && key < 120 + N_FRONT_CLASSES) {
    if (value < -1 || value > 2) throw std::runtime_error("set_tuning: class group launches take -1 (global), 0, 1 or 2");
    if (value != cls_gkern_[key - 120]) schedule_method_ = 0;
    cls_gkern_[key - 120] = value;
  } else if (key == 15) {
    if (value != 0 && value != 1) throw std::runtime_error("set_tuning: key 15 (FP32 Schur operand copies) is 0 or 1");
    if (value != schur_p32_) schedule_method_ = 0;
    schur_p32_ = value;
  } else if (key == 14) {
    if (value != 0 && value != 1) throw std::runtime_error("set_tuning: key 14 (small-front LU block steps) is 0 or 1");
    if (value != svb_small_) solve_method_ = 0;
    svb_small_ = value;
  } else if (key == 13) {
    if (value != 16 && value != 32 && value != 64)
      throw std::runtime_error("set_tuning: key 13 (LU row-kernel chunk) is 16, 32 or 64");
    if (value != lu_rows_w_) schedule_method_ = 0;
    lu_rows_w_ = value;
  } else if (key == 12) {
    if (value != 0 && value != 1) throw std::runtime_error("set_tuning: key 12 (LU lookahead) is 0 or 1");
    if (value != lu_lookahead_) schedule_method_ = 0;
    lu_lookahead_ = value;
  } else if (key == 11) {
    if (value != 0 && value != 1) throw std::runtime_error("set_tuning: key 11 (LU panel kernel) is 0 or 1");
    if (value != lu_panel_ll_) drop_graphs();  // captured launches name the kernel
    lu_panel_ll_ = value;
  } else if (key == 10) {
    if (value < 0) throw std::runtime_error("set_tuning: key 10 (low-precision Schur rows) takes a row count");
    if (value != lp_min_ss_) schedule_method_ = 0;
    lp_min_ss_ = value;
  } else if (key == 9) {
    if (value < 0) throw std::runtime_error("set_tuning: key 9 (fused small-front levels) takes 0 or a front count");
    if (value != small_fused_) schedule_method_ = 0;
    small_fused_ = value;
  } else if (key == 8) {
    if (value != 0 && value != 1) throw std::runtime_error("set_tuning: level solves are 0 or 1");
    if (value != solve_level_) solve_method_ = 0;
    solve_level_ = value;
  } else if (key == 7) {
    if (value < 0 || value > 2) throw std::runtime_error("set_tuning: panel-group launches are 0, 1 or 2");
    if (value != group_kernel_) schedule_method_ = 0;
    group_kernel_ = value;
  } else {
    throw std::runtime_error("set_tuning: unknown key (1: Cholesky panel group, 2: 128-tile threshold, "
                             "3: block-solve threshold, 4: whole-front LU limit, 5: Cholesky lookahead, "
                             "6: Cholesky diagonal blocks fused into the panel updates, 7: Cholesky panel groups "
                             "in one persistent launch, 8: Cholesky level solves in one launch per level, "
                             "9: fused small-front Cholesky levels, 10: smallest mean Schur block of a level "
                             "with low-precision products, 11: LU panel kernel (1 left-looking, 0 right-looking), 12: LU lookahead, 13: LU row-kernel chunk (16, 32, 64), 14: small-front LU block steps, 100 + c / 110 + c / 120 + c: panel group / tile "
                             "threshold / group launches of front size class c)");
  }
}

int FrontSolver::get_tuning(int key) const {
  if (key == 1) return chol_group();
  if (key == 2) return tile_threshold_;
  if (key == 3) return block_rhs_;
  if (key == 4) return front_small_;
  if (key == 5) return lookahead_;
  if (key == 6) return fuse_diag_;
  if (key == 7) return group_kernel_;
  if (key == 8) return solve_level_;
  if (key == 9) return small_fused_;
  if (key == 10) return lp_min_ss_;
  if (key == 11) return lu_panel_ll_;
  if (key == 12) return lu_lookahead_;
  if (key == 13) return lu_rows_w_;
  if (key == 14) return svb_small_;
  if (key == 15) return schur_p32_;
  if (key >= 100 && key < 100 + N_FRONT_CLASSES) return cls_group_[key - 100];
  if (key >= 110 && key < 110 + N_FRONT_CLASSES) return cls_tile_[key - 110];
  if (key >= 120 && key < 120 + N_FRONT_CLASSES) return cls_gkern_[key - 120];
  throw std::runtime_error("get_tuning: unknown key");
}

double FrontSolver::autotune(int method) {
  // Coordinate search over the schedule knobs on this pattern and these
  // values: the Cholesky panel group (panels per trailing update), the
  // mean tile extent from which a launch uses 128 x 128 GEMM tiles, the
  // Cholesky panel lookahead, then the per-size-class panel group and tile
  // threshold of every class the pattern has.  Each
  // candidate is timed as the better of two factorizations; the best
  // setting stays in place and its time is returned.
  auto time_it = [&]() {
    double best = 1e300;
    for (int r = 0; r < 2; ++r) {
      factor(method);
      best = std::min(best, t_factor_ms_);
    }
    return best;
  };
  const bool save_graph = use_graph_;
  use_graph_ = false;
  double best = time_it();
  if (method == UHM_CHOL) {
    int bg = chol_group();
    for (int g : {1, 2, 4, 6, 8, 12}) {
      if (g == bg) continue;
      set_chol_group(g);
      const double t = time_it();
      if (t < best) {
        best = t;
        bg = g;
      }
    }
    set_chol_group(bg);
  }
  int bt = tile_threshold_;
  for (int th : {512, 1024, 2048, 4096, 8192}) {
    if (th == bt) continue;
    set_tuning(2, th);
    const double t = time_it();
    if (t < best) {
      best = t;
      bt = th;
    }
  }
  set_tuning(2, bt);
  if (method == UHM_CHOL) {
    const int bl = lookahead_;
    set_tuning(5, 1 - bl);
    const double t = time_it();
    if (t < best) best = t;
    else set_tuning(5, bl);
  }
  // then per front size class present: its panel group (Cholesky) and its
  // tile threshold, the other classes held at their current values
  bool present[N_FRONT_CLASSES] = {false, false, false};
  for (const uhmk_front& F : fr_)
    if (F.fs > 0) present[front_class(F.fs)] = true;
  auto search = [&](int key, std::initializer_list<int> cand) {
    const int cur = get_tuning(key);
    int bv = cur;
    for (int v : cand) {
      if (v == cur) continue;
      set_tuning(key, v);
      const double t = time_it();
      if (t < best) {
        best = t;
        bv = v;
      }
    }
    set_tuning(key, bv);
  };
  for (int c = 0; c < N_FRONT_CLASSES; ++c) {
    if (!present[c]) continue;
    if (method == UHM_CHOL && c > 0) search(100 + c, {1, 2, 4, 6, 8, 12});
    search(110 + c, {512, 2048, 8192});
  }
  use_graph_ = save_graph;
  drop_graphs();
  return best;
}

void FrontSolver::set_reuse_source(std::unique_ptr<FrontSolver> prev) {
  reuse_.assign(fr_.size(), 0);
  reuse_src_.reset();
  if (!prev || !prev->factored_) return;
  for (int f = 0; f < (int)fr_.size(); ++f) reuse_[f] = order_[f]->reuse ? 1 : 0;
  reuse_src_ = std::move(prev);
  schedule_method_ = 0;
}

int FrontSolver::n_reused() const {
  int c = 0;
  for (char r : reuse_) c += r ? 1 : 0;
  return c;
}

void FrontSolver::adopt_reused() {
  // the previous pattern's front of the same element, its entries converted
  // to this front's storage when the precision of its level differs
  FrontSolver& P = *reuse_src_;
  std::unordered_map<const Element_*, int> old_front;
  for (int f = 0; f < (int)P.order_.size(); ++f) old_front[P.order_[f]] = f;
  for (int f = 0; f < (int)fr_.size(); ++f) {
    if (!reuse_[f]) continue;
    auto it = old_front.find(order_[f]);
    if (it == old_front.end() || P.fr_[it->second].n != fr_[f].n || P.fr_[it->second].fs != fr_[f].fs) {
      // not reusable after all: refactor it and everything above it
      for (int g = f; g >= 0; g = fr_[g].parent) reuse_[g] = 0;
    }
  }
  for (int f = 0; f < (int)fr_.size(); ++f) {
    if (!reuse_[f]) continue;
    const uhmk_front& O = P.fr_[old_front[order_[f]]];
    const uhmk_front& N = fr_[f];
    const long long cnt = (long long)N.n * N.n;
    const void* src = O.prec ? (const void*)(P.d_arena32_.as<float>() + O.off) : (const void*)(P.d_arena_.as<double>() + O.off);
    void* dst = N.prec ? (void*)(d_arena32_.as<float>() + N.off) : (void*)(d_arena_.as<double>() + N.off);
    check_k(uhmk_copy_front(src, O.prec, dst, N.prec, cnt, stream_), "reuse: front copy");
    if (N.fs)
      check_hip(hipMemcpyAsync(d_piv_.as<int>() + N.poff, P.d_piv_.as<int>() + O.poff, N.fs * sizeof(int),
                               hipMemcpyDeviceToDevice, stream_),
                "reuse: pivots");
  }
  check_hip(hipStreamSynchronize(stream_), "reuse");
  reuse_src_.reset();
}

int FrontSolver::factor(int method) {
  if (method != UHM_LU_PIV && method != UHM_CHOL)
    throw std::runtime_error("factor: only UHM_CHOL and UHM_LU_PIV are supported");
  if (!created_) create_buffers();
  if (reuse_src_) {
    adopt_reused();
    schedule_method_ = 0;
  }
  upload_stage();
  if (schedule_method_ != method) build_schedule(method);
  hipEvent_t e0, e1;
  check_hip(hipEventCreate(&e0), "event");
  check_hip(hipEventCreate(&e1), "event");
  check_hip(hipEventRecord(e0, stream_), "event");
  std::vector<hipEvent_t> pev;
  if (profile_) {
    stats_ = KernelStats();
    pev.resize(2 * launches_.size());
    for (auto& e : pev) check_hip(hipEventCreate(&e), "event");
    enqueue_factor(method, &pev);
  } else if (use_graph_) {
    // re-factorization on a fixed pattern: the launch sequence is captured
    // once and replayed
    if (!factor_graph_ || factor_graph_method_ != method) {
      if (factor_graph_) (void)hipGraphExecDestroy(factor_graph_);
      factor_graph_ = nullptr;
      factor_graph_ = capture([&] { enqueue_factor(method, nullptr); });
      factor_graph_method_ = method;
    }
    check_hip(hipGraphLaunch(factor_graph_, stream_), "graph launch");
  } else {
    enqueue_factor(method, nullptr);
  }
  check_hip(hipEventRecord(e1, stream_), "event");
  check_hip(hipEventSynchronize(e1), "factor");
  float ms = 0;
  check_hip(hipEventElapsedTime(&ms, e0, e1), "event");
  if (profile_) {
    for (size_t li = 0; li < launches_.size(); ++li) {
      float t = 0;
      check_hip(hipEventElapsedTime(&t, pev[2 * li], pev[2 * li + 1]), "event");
      const int k = (int)launches_[li].kind;
      stats_.ms[k] += t;
      stats_.flops[k] += launches_[li].flops;
      stats_.bytes[k] += launches_[li].bytes;
      stats_.launches[k] += 1;
    }
    for (auto& e : pev) (void)hipEventDestroy(e);
  }
  t_factor_ms_ = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::vector<int> hinfo(fr_.size());
  check_hip(hipMemcpy(hinfo.data(), d_info_.p, hinfo.size() * sizeof(int), hipMemcpyDeviceToHost), "info");
  int gave_up = 0;
  check_hip(hipMemcpy(&gave_up, d_flags_.p, sizeof(int), hipMemcpyDeviceToHost), "group flags");
  if (gave_up)
    throw std::runtime_error("factor: " + std::to_string(gave_up) +
                             " panel-group waits gave up (the factors are invalid)");
  factored_ = true;
  last_method_ = method;
  factor_nb_ = nb_;
  for (size_t f = 0; f < hinfo.size(); ++f)
    if (hinfo[f]) return order_[f]->id + 1;
  return 0;
}

bool FrontSolver::solve_schedule_current(int method) const {
  const int inv = (method == UHM_CHOL && factor_nb_ == 64) ? 1 : 0;
  const int blk = nrhs_ >= block_rhs_ ? 1 : 0;
  return solve_method_ == method && solve_inv_ == inv && solve_block_ == blk;
}

void FrontSolver::build_block_solve_schedule(int method) {
  // Block solves (many right-hand sides): per level, forward = gather,
  // 64-row block steps on ATL, one product xb -= ABL z_t; backward =
  // branch, one product z_t -= ATR xb / ABL^T xb, block steps.  Items
  // {f, first right-hand side of the 64-column tile, first row, 0}; the
  // tiles of one matrix chunk are adjacent so that they meet in L2.
  const bool chol = (method == UHM_CHOL);
  const int mode = chol ? (factor_nb_ == 64 ? 2 : 1) : 0;
  const bool small = mode == 0 && svb_small_;
  const int ntile = (nrhs_ + 63) / 64;
  auto step = [&](SolveStep::Kind kind, int off, int L) {
    SolveStep st;
    st.kind = kind;
    st.off = off;
    st.cnt = (int)sitems_.size() - off;
    st.prec = level_prec_[L];
    return st;
  };
  auto push = [&](int f, int r) {
    for (int t = 0; t < ntile; ++t) sitems_.push_back(uhmk_item{f, 64 * t, r, 0});
  };
  // LU: the first block step of the level's fronts with fs <= 32 (their
  // only one) by one wavefront per item (uhmk_solve_block_step_small)
  auto push_small = [&](int L, int fwd, std::vector<SolveStep>& list) {
    const int o = (int)sitems_.size();
    for (int f : levels_[L])
      if (fr_[f].fs > 0 && fr_[f].fs <= 32) push(f, 0);
    SolveStep st;
    st.kind = SolveStep::BSMALL;
    st.off = o;
    st.cnt = (int)sitems_.size() - o;
    st.prec = level_prec_[L];
    st.fwd = fwd;
    if (st.cnt) list.push_back(st);
  };
  // split-K for product launches with few items and long K (top levels):
  // enough workgroups to fill the chip, every slice at least 256 deep
  long long part_max = 0;
  auto split = [&](SolveStep& g, bool fwd) {
    double ksum = 0;
    for (int q = g.off; q < g.off + g.cnt; ++q) {
      const uhmk_front& fr = fr_[sitems_[q].f];
      ksum += fwd ? fr.fs : fr.n - fr.fs;
    }
    const double kmean = g.cnt ? ksum / g.cnt : 0.0;
    int ks = 1;
    if (g.cnt < 1024) ks = (int)std::min(8.0, std::ceil(1024.0 / g.cnt));
    ks = std::max(1, std::min(ks, (int)(kmean / 256.0)));
    g.ks = ks;
    if (ks > 1) part_max = std::max(part_max, (long long)g.cnt * ks * 4096);
  };
  const int nl = (int)levels_.size();
  std::vector<int> mfs(nl, 0);
  for (int L = 0; L < nl; ++L)
    for (int f : levels_[L]) mfs[L] = std::max(mfs[L], fr_[f].fs);
  // gather and branch: 64-row items, one wavefront per item and column group
  for (int L = 0; L < nl; ++L) {
    int o = (int)sitems_.size();
    for (int f : levels_[L])
      for (int r = 0; r < fr_[f].n; r += 64) sitems_.push_back(uhmk_item{f, 0, r, 0});
    SolveStep ga = step(SolveStep::GATHER, o, L);
    ga.cols = 1;
    sfwd_.push_back(ga);
    for (int k0 = 0; k0 < mfs[L]; k0 += 64) {
      if (k0 == 0 && small) push_small(L, 1, sfwd_);
      o = (int)sitems_.size();
      for (int f : levels_[L]) {
        const int fs = fr_[f].fs;
        if (fs <= k0 || (k0 == 0 && small && fs <= 32)) continue;
        push(f, k0);
        for (int r = k0 + 64; r < fs; r += 64) push(f, r);
      }
      SolveStep st = step(SolveStep::BSTEP, o, L);
      st.fwd = 1;
      st.k0 = k0;
      st.mode = mode;
      if (st.cnt) sfwd_.push_back(st);
    }
    o = (int)sitems_.size();
    for (int f : levels_[L]) {
      const int fs = fr_[f].fs, n = fr_[f].n;
      if (fs == 0) continue;
      for (int r = fs; r < n; r += 64) push(f, r);
    }
    SolveStep g = step(SolveStep::BGEMM, o, L);
    g.mode = 0;
    split(g, true);
    if (g.cnt) sfwd_.push_back(g);
  }
  for (int L = nl - 1; L >= 0; --L) {
    int o = (int)sitems_.size();
    for (int f : levels_[L])
      if (fr_[f].parent >= 0)
        for (int r = 0; r < fr_[f].n - fr_[f].fs; r += 64) sitems_.push_back(uhmk_item{f, 0, r, 0});
    SolveStep br = step(SolveStep::BRANCH, o, L);
    br.cols = 1;
    if (br.cnt) sbwd_.push_back(br);
    o = (int)sitems_.size();
    for (int f : levels_[L]) {
      const int fs = fr_[f].fs, n = fr_[f].n;
      if (fs == 0 || n == fs) continue;
      for (int r = 0; r < fs; r += 64) push(f, r);
    }
    SolveStep g = step(SolveStep::BGEMM, o, L);
    g.mode = chol ? 2 : 1;
    split(g, false);
    if (g.cnt) sbwd_.push_back(g);
    for (int k0 = ((mfs[L] - 1) / 64) * 64; k0 >= 0 && mfs[L] > 0; k0 -= 64) {
      o = (int)sitems_.size();
      for (int f : levels_[L]) {
        if (fr_[f].fs <= k0 || (k0 == 0 && small && fr_[f].fs <= 32)) continue;
        push(f, k0);
        for (int r = 0; r < k0; r += 64) push(f, r);
      }
      SolveStep st = step(SolveStep::BSTEP, o, L);
      st.fwd = 0;
      st.k0 = k0;
      st.mode = mode;
      if (st.cnt) sbwd_.push_back(st);
      if (k0 == 0 && small) push_small(L, 0, sbwd_);
    }
  }
  if (part_max > 0 && (size_t)part_max * sizeof(double) > d_part_.bytes) d_part_.alloc(part_max * sizeof(double));
}

void FrontSolver::build_solve_schedule(int method) {
  // Step lists of the level solves, built once per pattern and method.
  // Fronts with fs <= 1024 solve ATL in one workgroup; larger ones in
  // 512-row super-blocks whose couplings go through the gemv kernels, so
  // that a level's big fronts are read by many workgroups.  From block_rhs_
  // right-hand sides on, the block solves replace all of this.
  const bool chol = (method == UHM_CHOL);
  const bool blk = nrhs_ >= block_rhs_;
  const bool inv = chol && factor_nb_ == 64;
  sitems_.clear();
  sflist_.clear();
  sfwd_.clear();
  sbwd_.clear();
  auto rows = [&](int f, int a, int b, int cut, int step, int ca, int cb) {
    for (int r = a; r < std::min(b, cut); r += step) sitems_.push_back(uhmk_item{f, ca, r, cb});
    for (int r = std::max(a, cut); r < b; r += step) sitems_.push_back(uhmk_item{f, ca, r, cb});
  };
  auto step = [&](SolveStep::Kind kind, int off) {
    SolveStep st;
    st.kind = kind;
    st.off = off;
    st.cnt = (int)sitems_.size() - off;
    return st;
  };
  const int nl = (int)levels_.size();
  std::vector<int> tr_off(nl), tr_cnt(nl), mfs(nl), sbs(nl), nsb(nl);
  for (int L = 0; L < nl; ++L) {
    tr_off[L] = (int)sflist_.size();
    mfs[L] = 0;
    for (int f : levels_[L])
      if (fr_[f].fs > 0) {
        sflist_.push_back(f);
        mfs[L] = std::max(mfs[L], fr_[f].fs);
      }
    tr_cnt[L] = (int)sflist_.size() - tr_off[L];
    sbs[L] = mfs[L] <= 1024 ? 1024 : 512;
    nsb[L] = (mfs[L] + sbs[L] - 1) / sbs[L];
  }
  auto trsv = [&](int L, int sb) {
    SolveStep st;
    st.prec = level_prec_[L];
    st.kind = SolveStep::TRSV;
    st.off = tr_off[L];
    st.cnt = tr_cnt[L];
    st.max_fs = mfs[L];
    st.sb = sb;
    st.sb_size = sbs[L];
    return st;
  };
  auto rt_of = [&](int L) { return level_prec_[L] ? 32 : 16; };  // gemv row chunk: one 128-byte line
  auto cstep = [&](int L, int fwd, int k0, int off) {
    SolveStep st = step(SolveStep::CSTEP, off);
    st.fwd = fwd;
    st.k0 = k0;
    st.prec = level_prec_[L];
    return st;
  };
  if (blk) {
    build_block_solve_schedule(method);
  } else if (inv && solve_level_ && nrhs_ <= 8) {
    // one launch per level and direction (at most 8 right-hand sides per
    // launch); z marked first
    SolveStep mk;
    mk.kind = SolveStep::MARK;
    sfwd_.push_back(mk);
    for (int L = 0; L < nl; ++L) {
      int o = (int)sitems_.size();
      for (int f : levels_[L])
        for (int r = 0; r < fr_[f].n; r += 256) sitems_.push_back(uhmk_item{f, 0, r, 0});
      sfwd_.push_back(step(SolveStep::GATHER, o));
      o = (int)sitems_.size();
      for (int f : levels_[L])
        for (int r = 0; r < fr_[f].fs; r += 64) sitems_.push_back(uhmk_item{f, 0, r, 0});
      for (int f : levels_[L])
        if (fr_[f].fs > 0)
          for (int r = fr_[f].fs; r < fr_[f].n; r += 64) sitems_.push_back(uhmk_item{f, 1, r, 0});
      SolveStep lv = step(SolveStep::LEVEL, o);
      lv.fwd = 1;
      lv.prec = level_prec_[L];
      if (lv.cnt) sfwd_.push_back(lv);
    }
    for (int L = nl - 1; L >= 0; --L) {
      int o = (int)sitems_.size();
      for (int f : levels_[L])
        if (fr_[f].parent >= 0)
          for (int r = 0; r < fr_[f].n - fr_[f].fs; r += 256) sitems_.push_back(uhmk_item{f, 0, r, 0});
      SolveStep br = step(SolveStep::BRANCH, o);
      if (br.cnt) sbwd_.push_back(br);
      o = (int)sitems_.size();
      for (int f : levels_[L])
        for (int r = ((fr_[f].fs - 1) / 64) * 64; r >= 0 && fr_[f].fs > 0; r -= 64)
          sitems_.push_back(uhmk_item{f, 0, r, 0});
      SolveStep lv = step(SolveStep::LEVEL, o);
      lv.fwd = 0;
      lv.prec = level_prec_[L];
      if (lv.cnt) sbwd_.push_back(lv);
    }
  } else if (inv) {
    // Cholesky factors with diagonal-block inverses: 64-row block steps
    for (int L = 0; L < nl; ++L) {
      int o = (int)sitems_.size();
      for (int f : levels_[L])
        for (int r = 0; r < fr_[f].n; r += 256) sitems_.push_back(uhmk_item{f, 0, r, 0});
      sfwd_.push_back(step(SolveStep::GATHER, o));
      for (int k0 = 0; k0 < mfs[L]; k0 += 64) {
        o = (int)sitems_.size();
        for (int f : levels_[L]) {
          const int fs = fr_[f].fs, n = fr_[f].n;
          if (fs <= k0) continue;
          sitems_.push_back(uhmk_item{f, 0, k0, 0});
          for (int r = k0 + std::min(64, fs - k0); r < n; r += 64) sitems_.push_back(uhmk_item{f, 0, r, 0});
        }
        sfwd_.push_back(cstep(L, 1, k0, o));
      }
    }
    for (int L = nl - 1; L >= 0; --L) {
      int o = (int)sitems_.size();
      for (int f : levels_[L])
        if (fr_[f].parent >= 0)
          for (int r = 0; r < fr_[f].n - fr_[f].fs; r += 256) sitems_.push_back(uhmk_item{f, 0, r, 0});
      SolveStep br = step(SolveStep::BRANCH, o);
      if (br.cnt) sbwd_.push_back(br);
      o = (int)sitems_.size();
      for (int f : levels_[L]) {
        const int fs = fr_[f].fs, n = fr_[f].n;
        if (fs == 0 || n == fs) continue;
        for (int c0 = 0; c0 < fs; c0 += 16) sitems_.push_back(uhmk_item{f, fs, c0, n});
      }
      SolveStep g = step(SolveStep::GEMV, o);
      g.cols = 1;
      g.prec = level_prec_[L];
      if (g.cnt) sbwd_.push_back(g);
      for (int k0 = ((mfs[L] - 1) / 64) * 64; k0 >= 0 && mfs[L] > 0; k0 -= 64) {
        o = (int)sitems_.size();
        for (int f : levels_[L]) {
          if (fr_[f].fs <= k0) continue;
          sitems_.push_back(uhmk_item{f, 0, k0, 0});
          for (int c0 = 0; c0 < k0; c0 += 64) sitems_.push_back(uhmk_item{f, 0, c0, 0});
        }
        sbwd_.push_back(cstep(L, 0, k0, o));
      }
    }
  }
  for (int L = 0; L < nl && !inv && !blk; ++L) {
    const int rt = rt_of(L);
    int o = (int)sitems_.size();
    for (int f : levels_[L])
      for (int r = 0; r < fr_[f].n; r += 256) sitems_.push_back(uhmk_item{f, 0, r, 0});
    sfwd_.push_back(step(SolveStep::GATHER, o));
    for (int sb = 0; sb < nsb[L]; ++sb) {
      sfwd_.push_back(trsv(L, sb));
      sfwd_.back().fwd = 1;
      o = (int)sitems_.size();
      for (int f : levels_[L]) {
        const int fs = fr_[f].fs, n = fr_[f].n, s0 = sb * sbs[L];
        if (fs <= s0) continue;
        const int s1 = std::min(s0 + sbs[L], fs);
        rows(f, s1, n, fs, rt, s0, s1);
      }
      SolveStep g = step(SolveStep::GEMV, o);
      g.cols = 0;
      g.prec = level_prec_[L];
      if (g.cnt) sfwd_.push_back(g);
    }
  }
  for (int L = nl - 1; L >= 0 && !inv && !blk; --L) {
    const int rt = rt_of(L);
    int o = (int)sitems_.size();
    for (int f : levels_[L])
      if (fr_[f].parent >= 0)
        for (int r = 0; r < fr_[f].n - fr_[f].fs; r += 256) sitems_.push_back(uhmk_item{f, 0, r, 0});
    SolveStep br = step(SolveStep::BRANCH, o);
    if (br.cnt) sbwd_.push_back(br);
    o = (int)sitems_.size();
    for (int f : levels_[L]) {
      const int fs = fr_[f].fs, n = fr_[f].n;
      if (fs == 0 || n == fs) continue;
      if (chol)
        for (int c0 = 0; c0 < fs; c0 += 16) sitems_.push_back(uhmk_item{f, fs, c0, n});
      else
        rows(f, 0, fs, fs, rt, fs, n);
    }
    SolveStep g = step(SolveStep::GEMV, o);
    g.cols = chol ? 1 : 0;
    g.prec = level_prec_[L];
    if (g.cnt) sbwd_.push_back(g);
    for (int sb = nsb[L] - 1; sb >= 0; --sb) {
      sbwd_.push_back(trsv(L, sb));
      if (sb == 0) continue;
      o = (int)sitems_.size();
      for (int f : levels_[L]) {
        const int fs = fr_[f].fs, s0 = sb * sbs[L];
        if (fs <= s0) continue;
        const int s1 = std::min(s0 + sbs[L], fs);
        if (chol)
          for (int c0 = 0; c0 < s0; c0 += 16) sitems_.push_back(uhmk_item{f, s0, c0, s1});
        else
          rows(f, 0, s0, fs, rt, s0, s1);
      }
      SolveStep u = step(SolveStep::GEMV, o);
      u.cols = chol ? 1 : 0;
      u.prec = level_prec_[L];
      if (u.cnt) sbwd_.push_back(u);
    }
  }
  d_sitems_.upload(sitems_, stream_);
  d_sflist_.upload(sflist_, stream_);
  check_hip(hipStreamSynchronize(stream_), "solve schedule upload");
  for (auto& g : solve_graphs_)
    if (g.second) (void)hipGraphExecDestroy(g.second);
  solve_graphs_.clear();
  solve_method_ = method;
  solve_inv_ = inv ? 1 : 0;
  solve_block_ = blk ? 1 : 0;
}

void FrontSolver::run_solve_steps(const std::vector<SolveStep>& steps, bool chol, const double* rhs) {
  const uhmk_front* F = dF_.as<uhmk_front>();
  const uhmk_item* SI = d_sitems_.as<uhmk_item>();
  const int* SF = d_sflist_.as<int>();
  double* X = d_x_.as<double>();
  for (const SolveStep& st : steps) {
    const void* A = arena_ptr(st.prec);
    int rc = 0;
    switch (st.kind) {
      case SolveStep::GATHER:
rc = (st.cols ? uhmk_solve_gather_cols : uhmk_solve_gather)(
            F, SI + st.off, st.cnt, chol ? 0 : 1, d_srcbeg_.as<int>(), d_srcs_.as<uhmk_source>(), d_inv_.as<int>(),
            d_perm_.as<int>(), X, rhs, nrhs_, stream_);
        break;
      case SolveStep::TRSV:
        rc = uhmk_solve_trsv(F, SF + st.off, st.cnt, st.max_fs, chol ? (factor_nb_ == 64 ? 2 : 1) : 0, st.fwd, st.sb,
                             st.sb_size, A, st.prec, X, nrhs_, stream_);
        break;
      case SolveStep::GEMV:
        rc = uhmk_solve_gemv(F, SI + st.off, st.cnt, st.cols, A, st.prec, X, solve_inv_ ? d_z_.as<double>() : X,
                             nrhs_, stream_);
        break;
      case SolveStep::CSTEP:
        rc = uhmk_solve_chol_step(F, SI + st.off, st.cnt, st.fwd, st.k0, A, st.prec, X, d_z_.as<double>(), nrhs_,
                                  stream_);
        break;
      case SolveStep::BRANCH:
        rc = (st.cols ? uhmk_solve_branch_cols : uhmk_solve_branch)(F, SI + st.off, st.cnt, d_emap_.as<int>(), X,
                                                                    nrhs_, stream_);
        break;
      case SolveStep::BSMALL:
        rc = uhmk_solve_block_step_small(F, SI + st.off, st.cnt, st.fwd, A, st.prec, X, d_z_.as<double>(), nrhs_,
                                         stream_);
        break;
      case SolveStep::BSTEP:
        rc = uhmk_solve_block_step(F, SI + st.off, st.cnt, st.mode, st.fwd, st.k0, A, st.prec, X, d_z_.as<double>(),
                                   nrhs_, stream_);
        break;
      case SolveStep::LEVEL:
        rc = uhmk_solve_chol_level(F, SI + st.off, st.cnt, st.fwd, A, st.prec, X, d_z_.as<double>(), nrhs_,
                                   d_serr_.as<int>(), stream_);
        break;
      case SolveStep::MARK:
        rc = uhmk_solve_mark(d_z_.as<double>(), x_elems_, stream_);
        if (!rc) rc = uhmk_fill_int(d_serr_.as<int>(), 1, 0, stream_);
        break;
      case SolveStep::BGEMM:
        rc = uhmk_solve_block_gemm(F, SI + st.off, st.cnt, st.mode, A, st.prec, X, d_z_.as<double>(), nrhs_, st.ks,
                                   d_part_.as<double>(), stream_);
        break;
    }
    check_k(rc, "solve");
  }
}

void FrontSolver::run_solve(bool chol, const double* rhs) {
  if (!use_graph_) {
    run_solve_steps(sfwd_, chol, rhs);
    run_solve_steps(sbwd_, chol, rhs);
    return;
  }
  hipGraphExec_t& g = solve_graphs_[rhs];
  if (!g)
    g = capture([&] {
      run_solve_steps(sfwd_, chol, rhs);
      run_solve_steps(sbwd_, chol, rhs);
    });
  check_hip(hipGraphLaunch(g, stream_), "graph launch");
}

void FrontSolver::solve(int method) {
  if (!factored_) throw std::runtime_error("solve: not factorized");
  if (method != last_method_) throw std::runtime_error("solve: method differs from the factorization's");
  upload_rhs();
  if (!solve_schedule_current(method))
    build_solve_schedule(method);
  const bool chol = (method == UHM_CHOL);
  hipEvent_t e0, e1;
  check_hip(hipEventCreate(&e0), "event");
  check_hip(hipEventCreate(&e1), "event");
  check_hip(hipEventRecord(e0, stream_), "event");
  run_solve(chol, d_b_.as<double>());
  check_hip(hipEventRecord(e1, stream_), "event");
  check_hip(hipEventSynchronize(e1), "solve");
  float ms = 0;
  check_hip(hipEventElapsedTime(&ms, e0, e1), "event");
  t_solve_ms_ = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  check_solve_err();
}

void FrontSolver::check_solve_err() {
  int e = 0;
  check_hip(hipMemcpy(&e, d_serr_.p, sizeof(int), hipMemcpyDeviceToHost), "solve error word");
  if (e) throw std::runtime_error("solve: " + std::to_string(e) + " level-solve waits gave up (x is invalid)");
}

// ------------------------------------------------------------ assembled A
void FrontSolver::build_csr() {
  // Assembled A = sum over leaves of P_e K_e P_e^T in the global numbering,
  // the DSparse_::import_element semantics (UHM/src/interf/sparse.cxx:196-267,
  // axpy = 1).  Symbolic on the host; values summed on the device from the
  // leaf stages in leaf order, so re-generated values need no host work.
  const int nf = (int)fr_.size();
  std::vector<std::vector<long long>> gidx(nf);
  std::vector<int> leaves;
  for (int f = 0; f < nf; ++f) {
    if (fr_[f].soff < 0) continue;
    leaves.push_back(f);
    Element_* e = order_[f];
    auto& G = gidx[f];
    G.resize(fr_[f].n);
    for (int i = 0; i < e->fs; ++i) G[i] = goff_[f] + i;
    for (auto& p : e->schur) {
      if (p.first->offset < 0) throw std::runtime_error("residual: unnumbered (boundary) nodes are not supported");
      for (int d = 0; d < p.first->n_dof; ++d) G[e->fs + p.second + d] = p.first->offset + d;
    }
  }
  const long long N = ndof_;
  // rows -> (leaf, local index)
  std::vector<long long> rcnt(N + 1, 0);
  for (int f : leaves)
    for (long long gi : gidx[f]) rcnt[gi + 1]++;
  for (long long i = 0; i < N; ++i) rcnt[i + 1] += rcnt[i];
  std::vector<std::pair<int, int>> rl(rcnt[N]);
  {
    std::vector<long long> pos(rcnt.begin(), rcnt.end() - 1);
    for (int f : leaves)
      for (int i = 0; i < fr_[f].n; ++i) rl[pos[gidx[f][i]]++] = {f, i};
  }
  // compact leaf layout of the leaf-product residual
  std::vector<long long> loff(nf, -1);
  long long NL = 0;
  for (int f : leaves) {
    loff[f] = NL;
    NL += fr_[f].n;
  }
  if (NL >= (1LL << 31)) throw std::runtime_error("residual: leaf rows exceed the 32-bit item range");
  std::vector<long long> rowptr(N + 1, 0);
  std::vector<int> colidx;
  std::vector<long long> lpos;
  std::vector<long long> cptr(1, 0), cidx;
  std::vector<long long> bptr(N + 1, 0), bidx, bn;
  std::vector<int> mark(N, -1);
  std::vector<long long> cols;
  std::vector<std::vector<long long>> slot_contrib;
  colidx.reserve(rl.size() * 8);
  for (long long g = 0; g < N; ++g) {
    cols.clear();
    for (long long q = rcnt[g]; q < rcnt[g + 1]; ++q) {
      const int f = rl[q].first;
      for (long long c : gidx[f])
        if (mark[c] != (int)(g & 0x7fffffff)) {
          mark[c] = (int)(g & 0x7fffffff);
          cols.push_back(c);
        }
    }
    std::sort(cols.begin(), cols.end());
    for (size_t k = 0; k < cols.size(); ++k) {
      colidx.push_back((int)cols[k]);
      mark[cols[k]] = -1;
    }
    rowptr[g + 1] = (long long)colidx.size();
    // contributions per slot, leaf order (rl is in leaf order within a row)
    auto slot_of = [&](long long c) {
      return (long long)(std::lower_bound(cols.begin(), cols.end(), c) - cols.begin());
    };
    std::vector<std::pair<long long, long long>> tmp;
    for (long long q = rcnt[g]; q < rcnt[g + 1]; ++q) {
      const int f = rl[q].first, i = rl[q].second;
      const int n = fr_[f].n;
      for (int j = 0; j < n; ++j) tmp.push_back({slot_of(gidx[f][j]), fr_[f].soff + i + (long long)j * n});
      bidx.push_back(fr_[f].xoff + i);
      bn.push_back(n);
      lpos.push_back(loff[f] * nrhs_ + i);
    }
    bptr[g + 1] = (long long)bidx.size();
    std::stable_sort(tmp.begin(), tmp.end(),
                     [](const std::pair<long long, long long>& a, const std::pair<long long, long long>& b) {
                       return a.first < b.first;
                     });
    size_t t = 0;
    for (size_t k = 0; k < cols.size(); ++k) {
      while (t < tmp.size() && tmp[t].first == (long long)k) cidx.push_back(tmp[t++].second);
      cptr.push_back((long long)cidx.size());
    }
  }
  nnz_ = (long long)colidx.size();
  d_rowptr_.upload(rowptr, stream_);
  d_colidx_.upload(colidx, stream_);
  d_cptr_.upload(cptr, stream_);
  d_cidx_.upload(cidx, stream_);
  d_bptr_.upload(bptr, stream_);
  d_bidx_.upload(bidx, stream_);
  std::vector<int> bn32(bn.begin(), bn.end());
  d_bn_.upload(bn32, stream_);
  {
    std::vector<long long> lg(std::max(1LL, NL));
    std::vector<uhmk_item> gi, mi;
    for (int f : leaves) {
      const int n = fr_[f].n, lo = (int)loff[f];
      for (int i = 0; i < n; ++i) lg[loff[f] + i] = gidx[f][i];
      for (int r = 0; r < n; r += 256) gi.push_back(uhmk_item{f, 0, r, lo});
      for (int r = 0; r < n; r += 64)
        for (int c0 = 0; c0 < nrhs_; c0 += 64) mi.push_back(uhmk_item{f, c0, r, lo});
    }
    d_lg_.upload(lg, stream_);
    d_lpos_.upload(lpos, stream_);
    d_lgi_.upload(gi, stream_);
    d_lmi_.upload(mi, stream_);
    n_lgi_ = (int)gi.size();
    n_lmi_ = (int)mi.size();
    d_xl_.alloc(std::max(1LL, NL * nrhs_) * sizeof(double));
    d_yl_.alloc(std::max(1LL, NL * nrhs_) * sizeof(double));
    d_amax_.alloc(sizeof(double));
  }
  d_val_.alloc(std::max(1LL, nnz_) * sizeof(double));
  d_xg_.alloc(std::max(1LL, N * nrhs_) * sizeof(double));
  d_bg_.alloc(std::max(1LL, N * nrhs_) * sizeof(double));
  d_r_.alloc(std::max(1LL, N * nrhs_) * sizeof(double));
  d_norms_.alloc(4 * sizeof(double) * std::max(1, nrhs_));
  check_hip(hipStreamSynchronize(stream_), "build_csr");
  csr_built_ = true;
  vals_dirty_ = bg_dirty_ = true;
  refresh_csr();
}

void FrontSolver::refresh_csr() {
  // Values and the assembled right-hand side follow the current inputs; the
  // pattern is built once.
  if (!csr_built_) {
    build_csr();
    return;
  }
  upload_stage();
  upload_rhs();
  const long long N = ndof_;
  if (vals_dirty_) {
    check_k(uhmk_csr_values(nnz_, d_cptr_.as<long long>(), d_cidx_.as<long long>(), d_stage_.as<double>(),
                            d_val_.as<double>(), stream_),
            "csr values");
    check_hip(hipMemsetAsync(d_amax_.p, 0, sizeof(double), stream_), "memset amax");
    check_k(uhmk_rowsum_max(N, d_rowptr_.as<long long>(), d_val_.as<double>(), d_amax_.as<double>(), stream_),
            "row sums");
    check_hip(hipMemcpyAsync(&amax_, d_amax_.p, sizeof(double), hipMemcpyDeviceToHost, stream_), "amax");
    check_hip(hipStreamSynchronize(stream_), "amax");
  }
  if (bg_dirty_)
    for (int col = 0; col < nrhs_; ++col)
      check_k(uhmk_gather_rhs(N, d_bptr_.as<long long>(), d_bidx_.as<long long>(), d_bn_.as<int>(), d_b_.as<double>(),
                              d_bg_.as<double>() + (long long)col * N, col, stream_),
              "gather rhs");
  vals_dirty_ = bg_dirty_ = false;
}

void FrontSolver::residual_all(std::vector<double>& h) {
  // r = b - A x for every column from the global x (leaf gather, leaf
  // products, row gather); h[4 c ..] = {max|r|, max row sum |A|, max|x|,
  // max|b|} of column c.
  const long long N = ndof_;
  const uhmk_front* F = dF_.as<uhmk_front>();
  double* nrm = d_norms_.as<double>();
  h.assign(4 * (size_t)nrhs_, 0.0);
  check_hip(hipMemsetAsync(nrm, 0, 4 * sizeof(double) * nrhs_, stream_), "memset norms");
  check_k(uhmk_leaf_gather(F, d_lgi_.as<uhmk_item>(), n_lgi_, d_lg_.as<long long>(), d_xg_.as<double>(), N,
                           d_xl_.as<double>(), nrhs_, stream_),
          "leaf gather");
  check_k(uhmk_leaf_mul(F, d_lmi_.as<uhmk_item>(), n_lmi_, d_stage_.as<double>(), d_xl_.as<double>(),
                        d_yl_.as<double>(), nrhs_, stream_),
          "leaf products");
  check_k(uhmk_leaf_residual(N, d_bptr_.as<long long>(), d_lpos_.as<long long>(), d_bn_.as<int>(),
                             d_yl_.as<double>(), d_xg_.as<double>(), d_bg_.as<double>(), d_r_.as<double>(), nrhs_,
                             nrm, stream_),
          "residual");
  check_hip(hipMemcpyAsync(h.data(), nrm, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream_), "norms");
  check_hip(hipStreamSynchronize(stream_), "residual");
  for (int c = 0; c < nrhs_; ++c) h[4 * c + 1] = amax_;
}

void FrontSolver::set_refinement(int kind, int restart, double inner_tol) {
  if (kind < 0 || kind > 2) throw std::runtime_error("set_refinement: kind 0 (classical), 1 (GMRES-IR) or 2 (auto)");
  if (restart < 1 || restart > 31) throw std::runtime_error("set_refinement: GMRES restart length must be 1..31");
  if (!(inner_tol > 0.0 && inner_tol < 1.0)) throw std::runtime_error("set_refinement: inner tolerance in (0, 1)");
  refine_kind_ = kind;
  gmres_m_ = restart;
  gmres_tol_ = inner_tol;
}

void FrontSolver::apply_matrix(const double* v, double* out) {
  const uhmk_front* F = dF_.as<uhmk_front>();
  check_k(uhmk_leaf_gather(F, d_lgi_.as<uhmk_item>(), n_lgi_, d_lg_.as<long long>(), v, ndof_, d_xl_.as<double>(),
                           nrhs_, stream_),
          "leaf gather");
  check_k(uhmk_leaf_mul(F, d_lmi_.as<uhmk_item>(), n_lmi_, d_stage_.as<double>(), d_xl_.as<double>(),
                        d_yl_.as<double>(), nrhs_, stream_),
          "leaf products");
  check_k(uhmk_leaf_residual(ndof_, d_bptr_.as<long long>(), d_lpos_.as<long long>(), d_bn_.as<int>(),
                             d_yl_.as<double>(), nullptr, nullptr, out, nrhs_, nullptr, stream_),
          "product");
}

void FrontSolver::apply_factors(bool chol, const double* r, double* out) {
  // r on the factor rows of the owning fronts (the schur rows of the
  // refinement buffer stay zero), the level solves, the factor rows back
  const uhmk_front* F = dF_.as<uhmk_front>();
  const int nf = (int)fr_.size();
  check_k(uhmk_front_global_all(F, nf, d_goff_.as<long long>(), d_rf_.as<double>(), const_cast<double*>(r), nrhs_,
                                ndof_, 1, stream_),
          "scatter r");
  run_solve(chol, d_rf_.as<double>());
  if (out)
    check_k(uhmk_front_global_all(F, nf, d_goff_.as<long long>(), d_x_.as<double>(), out, nrhs_, ndof_, 0, stream_),
            "gather d");
  ++refine_solves_;
}

void FrontSolver::gmres_correction(bool chol) {
  // One GMRES-IR correction (Carson and Higham, SIAM J. Sci. Comput. 40
  // (2018) A817): GMRES(m) on the left-preconditioned system
  // M^-1 A d = M^-1 r from d = 0, every right-hand side in lockstep,
  // classical Gram-Schmidt with one reorthogonalisation, Givens rotations on
  // the host; each column stops at ||M^-1 (r - A d)|| <= inner_tol
  // ||M^-1 r|| (or at a happy breakdown); then x += V y.  M^-1 and A are
  // applied in FP64 arithmetic (factors stored in their level precision);
  // r is the current residual in d_r_.
  const long long N = ndof_;
  const int nr = nrhs_, m = gmres_m_;
  const long long vs = N * nr;
  // grow-only: the basis and the reduction buffers are sized once per
  // pattern and restart length, not reallocated per correction
  d_V_.reserve((size_t)(m + 1) * vs * sizeof(double));
  d_dpart_.reserve(std::max(1LL, uhmk_block_dots_scratch(N, nr)) * sizeof(double));
  d_dout_.reserve((size_t)(m + 1) * nr * sizeof(double));
  d_coef_.reserve((size_t)(m + 1) * nr * sizeof(double));
  double* V = d_V_.as<double>();
  double* dout = d_dout_.as<double>();
  double* coef = d_coef_.as<double>();
  std::vector<double> hv((size_t)(m + 1) * nr), sc(nr), beta(nr);
  std::vector<double> H((size_t)nr * (m + 1) * m, 0.0), cs((size_t)nr * m), sn((size_t)nr * m);
  std::vector<double> g((size_t)nr * (m + 1), 0.0);
  std::vector<int> kc(nr, -1);  // steps at which a column stopped (-1: running)
  auto Hc = [&](int c, int i, int j) -> double& { return H[((size_t)c * (m + 1) + i) * m + j]; };
  auto fetch = [&](int k) {
    check_hip(hipMemcpyAsync(hv.data(), dout, (size_t)k * nr * sizeof(double), hipMemcpyDeviceToHost, stream_),
              "gmres dots");
    check_hip(hipStreamSynchronize(stream_), "gmres dots");
  };
  auto scale = [&](double* w) {
    check_hip(hipMemcpyAsync(coef, sc.data(), nr * sizeof(double), hipMemcpyHostToDevice, stream_), "gmres scale");
    check_k(uhmk_scale_columns(w, N, nr, coef, stream_), "gmres scale");
  };
  auto norms = [&](double* w) {
    check_k(uhmk_block_dots(w, vs, 1, w, N, nr, d_dpart_.as<double>(), dout, stream_), "gmres norm");
    fetch(1);
  };
  // v0 = M^-1 r / beta
  apply_factors(chol, d_r_.as<double>(), V);
  norms(V);
  for (int c = 0; c < nr; ++c) {
    beta[c] = std::sqrt(std::max(0.0, hv[c]));
    g[(size_t)c * (m + 1)] = beta[c];
    if (!(beta[c] > 0.0)) kc[c] = 0;
    sc[c] = beta[c] > 0.0 ? 1.0 / beta[c] : 0.0;
  }
  scale(V);
  int k = 0;
  std::vector<double> hj((size_t)(m + 1) * nr);
  for (int j = 0; j < m; ++j) {
    bool any = false;
    for (int c = 0; c < nr; ++c) any |= kc[c] < 0;
    if (!any) break;
    double* W = V + (long long)(j + 1) * vs;
    apply_matrix(V + (long long)j * vs, d_r_.as<double>());
    apply_factors(chol, d_r_.as<double>(), W);
    std::fill(hj.begin(), hj.end(), 0.0);
    for (int pass = 0; pass < 2; ++pass) {
      check_k(uhmk_block_dots(V, vs, j + 1, W, N, nr, d_dpart_.as<double>(), dout, stream_), "gmres dots");
      check_k(uhmk_basis_combine(V, vs, j + 1, dout, -1.0, 1.0, W, N, nr, stream_), "gmres orthogonalise");
      fetch(j + 1);
      for (size_t q = 0; q < (size_t)(j + 1) * nr; ++q) hj[q] += hv[q];
    }
    norms(W);
    for (int c = 0; c < nr; ++c) {
      sc[c] = 0.0;
      if (kc[c] >= 0) continue;
      for (int i = 0; i <= j; ++i) Hc(c, i, j) = hj[(size_t)i * nr + c];
      const double hn = std::sqrt(std::max(0.0, hv[c]));
      Hc(c, j + 1, j) = hn;
      for (int i = 0; i < j; ++i) {
        const double a = Hc(c, i, j), b = Hc(c, i + 1, j);
        const double ci = cs[(size_t)c * m + i], si = sn[(size_t)c * m + i];
        Hc(c, i, j) = ci * a + si * b;
        Hc(c, i + 1, j) = -si * a + ci * b;
      }
      const double a = Hc(c, j, j), b = Hc(c, j + 1, j);
      const double rr = std::hypot(a, b);
      const double cj = rr > 0.0 ? a / rr : 1.0, sj = rr > 0.0 ? b / rr : 0.0;
      cs[(size_t)c * m + j] = cj;
      sn[(size_t)c * m + j] = sj;
      Hc(c, j, j) = rr;
      Hc(c, j + 1, j) = 0.0;
      double* gc = &g[(size_t)c * (m + 1)];
      gc[j + 1] = -sj * gc[j];
      gc[j] = cj * gc[j];
      if (std::fabs(gc[j + 1]) <= gmres_tol_ * beta[c] || !(hn > 0.0)) kc[c] = j + 1;
      else sc[c] = 1.0 / hn;
    }
    scale(W);
    k = j + 1;
  }
  // y = R^-1 g per column; x += V y
  std::vector<double> y((size_t)(m + 1) * nr, 0.0);
  for (int c = 0; c < nr; ++c) {
    const int kk = kc[c] < 0 ? k : kc[c];
    const double* gc = &g[(size_t)c * (m + 1)];
    for (int i = kk - 1; i >= 0; --i) {
      double t = gc[i];
      for (int l = i + 1; l < kk; ++l) t -= Hc(c, i, l) * y[(size_t)l * nr + c];
      const double d = Hc(c, i, i);
      y[(size_t)i * nr + c] = d != 0.0 ? t / d : 0.0;
    }
  }
  if (k > 0) {
    check_hip(hipMemcpyAsync(coef, y.data(), (size_t)k * nr * sizeof(double), hipMemcpyHostToDevice, stream_),
              "gmres y");
    check_k(uhmk_basis_combine(V, vs, k, coef, 1.0, 1.0, d_xg_.as<double>(), N, nr, stream_), "x += V y");
  }
  check_hip(hipStreamSynchronize(stream_), "gmres");
}

int FrontSolver::refine(int method, double tol, int max_iter) {
  // Iterative refinement with FP64 residuals on the assembled matrix and
  // corrections through the (possibly low-precision) factors: the slot of
  // Mesh_::improve_{chol,lu_piv} (UHM/src/mesh/chol.cxx:136-149,
  // UHM/src/mesh/lu_piv.cxx:232-245), whose reference body is a stub.
  if (!factored_) throw std::runtime_error("refine: not factorized");
  if (method != last_method_) throw std::runtime_error("refine: method differs from the factorization's");
  refresh_csr();
  if (!solve_schedule_current(method))
    build_solve_schedule(method);
  const bool chol = (method == UHM_CHOL);
  const long long N = ndof_;
  const uhmk_front* F = dF_.as<uhmk_front>();
  const int nf = (int)fr_.size();
  const auto t0 = std::chrono::steady_clock::now();
  refine_hist_.clear();
  refine_solves_ = 0;
  const long long* GO = d_goff_.as<long long>();
  check_k(uhmk_front_global_all(F, nf, GO, d_x_.as<double>(), d_xg_.as<double>(), nrhs_, N, 0, stream_), "gather x");
  std::vector<double> h;
  auto worst_eta = [&]() {
    residual_all(h);
    double worst = 0.0;
    for (int col = 0; col < nrhs_; ++col) {
      const double* hc = &h[4 * col];
      const double e = hc[0] / (hc[1] * hc[2] + hc[3]);
      if (!(e == e)) return e;
      worst = std::max(worst, e);
    }
    return worst;
  };
  double eta = worst_eta();
  refine_hist_.push_back(eta);
  bool krylov = refine_kind_ == 1;
  int it = 0;
  while (!(eta <= tol) && it < max_iter && eta == eta) {
    if (krylov) {
      gmres_correction(chol);
    } else {
      // d = M^-1 r (front layout, in d_x_), x += d
      apply_factors(chol, d_r_.as<double>(), nullptr);
      check_k(uhmk_front_global_all(F, nf, GO, d_x_.as<double>(), d_xg_.as<double>(), nrhs_, N, 2, stream_),
              "x += d");
    }
    ++it;
    const double e2 = worst_eta();
    refine_hist_.push_back(e2);
    if (refine_kind_ == 2 && !krylov && !(e2 <= 0.1 * eta)) krylov = true;
    eta = e2;
  }
  eta_ = eta;
  if (it > 0) {
    // the refined solution back into the fronts (factor rows, then the
    // branch pass for the schur rows) so that copy_out sees it; without a
    // correction the fronts already hold it
    check_k(uhmk_front_global_all(F, nf, GO, d_x_.as<double>(), d_xg_.as<double>(), nrhs_, N, 1, stream_),
            "scatter x");
    std::vector<SolveStep> br;
    for (const SolveStep& st : sbwd_)
      if (st.kind == SolveStep::BRANCH) br.push_back(st);
    run_solve_steps(br, chol, d_b_.as<double>());
  }
  check_hip(hipStreamSynchronize(stream_), "refine");
  t_refine_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  check_solve_err();
  return it;
}

void FrontSolver::check(int) {
  // A correct replacement for check_{chol,lu_piv} + get_residual: the true
  // residual r = b - A x against the assembled matrix (the reference's
  // hierarchical check multiplies through a no-op trmm,
  // LINAL/src/dense/trmm.cxx:104-130).
  if (!factored_) throw std::runtime_error("check: not factorized");
  refresh_csr();
  const long long N = ndof_;
  check_k(uhmk_front_global_all(dF_.as<uhmk_front>(), (int)fr_.size(), d_goff_.as<long long>(), d_x_.as<double>(),
                                d_xg_.as<double>(), nrhs_, N, 0, stream_),
          "gather x");
  std::vector<double> h;
  residual_all(h);
  std::vector<double> r((size_t)N * nrhs_);
  check_hip(hipMemcpy(r.data(), d_r_.p, r.size() * sizeof(double), hipMemcpyDeviceToHost), "residual");
  eta_ = 0.0;
  residual_1norm_ = 0.0;
  for (double& v : norms_) v = 0.0;
  for (int col = 0; col < nrhs_; ++col) {
    const double* hc = &h[4 * col];
    for (int k = 0; k < 4; ++k) norms_[k] = std::max(norms_[k], hc[k]);
    const double eta = hc[0] / (hc[1] * hc[2] + hc[3]);
    eta_ = (eta == eta) ? std::max(eta_, eta) : eta;
  }
  for (double v : r) residual_1norm_ += std::fabs(v);
}

void FrontSolver::csr(std::vector<long long>& rowptr, std::vector<int>& col, std::vector<double>& val) {
  refresh_csr();
  rowptr.resize(ndof_ + 1);
  col.resize(nnz_);
  val.resize(nnz_);
  check_hip(hipMemcpy(rowptr.data(), d_rowptr_.p, rowptr.size() * sizeof(long long), hipMemcpyDeviceToHost), "csr");
  check_hip(hipMemcpy(col.data(), d_colidx_.p, col.size() * sizeof(int), hipMemcpyDeviceToHost), "csr");
  check_hip(hipMemcpy(val.data(), d_val_.p, val.size() * sizeof(double), hipMemcpyDeviceToHost), "csr");
}

// ------------------------------------------------------------ results
void FrontSolver::fetch_x(int f, int m, int nr, const int* map, double* out) {
  if (nr != nrhs_) throw std::runtime_error("copy_out: rhs column count differs from n_rhs");
  const int n = fr_[f].n;
  std::vector<double> h((size_t)n * nrhs_);
  check_hip(hipMemcpy(h.data(), d_x_.as<double>() + fr_[f].xoff, h.size() * sizeof(double), hipMemcpyDeviceToHost),
            "fetch x");
  for (int r = 0; r < nr; ++r)
    for (int i = 0; i < m; ++i) out[i + (size_t)r * m] = h[map[i] + (size_t)r * n];
}

void FrontSolver::fetch_front_raw(int f, double* out) {
  const size_t nn = (size_t)fr_[f].n * fr_[f].n;
  if (fr_[f].prec) {
    std::vector<float> h(nn);
    check_hip(hipMemcpy(h.data(), d_arena32_.as<float>() + fr_[f].off, nn * sizeof(float), hipMemcpyDeviceToHost),
              "fetch front");
    for (size_t k = 0; k < nn; ++k) out[k] = h[k];
  } else {
    check_hip(hipMemcpy(out, d_arena_.as<double>() + fr_[f].off, nn * sizeof(double), hipMemcpyDeviceToHost),
              "fetch front");
  }
}

void FrontSolver::fetch_front(int f, int m, const int* map, double* out) {
  const int n = fr_[f].n;
  std::vector<double> h((size_t)n * n);
  fetch_front_raw(f, h.data());
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) out[i + (size_t)j * m] = h[map[i] + (size_t)map[j] * n];
}

void FrontSolver::fetch_pivots(int f, int* out) {
  if (!fr_[f].fs) return;
  check_hip(hipMemcpy(out, d_piv_.as<int>() + fr_[f].poff, fr_[f].fs * sizeof(int), hipMemcpyDeviceToHost),
            "fetch pivots");
}

void FrontSolver::fetch_stage(int f, double* out) {
  const int n = fr_[f].n;
  if (fr_[f].soff < 0) {
    std::fill(out, out + (size_t)n * n, 0.0);
    return;
  }
  if (stage_dirty_) upload_stage();
  check_hip(hipMemcpy(out, d_stage_.as<double>() + fr_[f].soff, (size_t)n * n * sizeof(double),
                      hipMemcpyDeviceToHost),
            "fetch stage");
}

void FrontSolver::fetch_rhs_front(int f, double* out) {
  if (rhs_dirty_) upload_rhs();
  const size_t len = (size_t)fr_[f].n * nrhs_;
  check_hip(hipMemcpy(out, d_b_.as<double>() + fr_[f].xoff, len * sizeof(double), hipMemcpyDeviceToHost),
            "fetch rhs");
}

void FrontSolver::fetch_x_global(double* out) {
  if (!d_xg_.p) d_xg_.alloc(std::max(1LL, ndof_ * nrhs_) * sizeof(double));
  check_k(uhmk_front_global_all(dF_.as<uhmk_front>(), (int)fr_.size(), d_goff_.as<long long>(), d_x_.as<double>(),
                                d_xg_.as<double>(), nrhs_, ndof_, 0, stream_),
          "gather x");
  check_hip(hipStreamSynchronize(stream_), "gather x");
  check_hip(hipMemcpy(out, d_xg_.p, ndof_ * nrhs_ * sizeof(double), hipMemcpyDeviceToHost), "fetch x global");
}

void FrontSolver::fetch_b_global(double* out) {
  refresh_csr();
  check_hip(hipMemcpy(out, d_bg_.p, ndof_ * nrhs_ * sizeof(double), hipMemcpyDeviceToHost), "fetch b global");
}

}  // namespace uhm
