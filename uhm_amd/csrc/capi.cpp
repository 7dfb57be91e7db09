// extern "C" entry points: the reference's Fortran/C interface names
// (UHM/include/uhm/wrapper/fort.hxx:47-190) over the host mesh API and the
// device FrontSolver, plus the extensions declared in include/uhm_amd.h.
// Every call catches exceptions and records the message for
// uhm_get_last_error_ instead of aborting.
#include <sys/time.h>

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "mesh.hpp"
#include "solver.hpp"
#include "uhm_amd.h"

namespace uhm {
void build_synthetic_mesh(Mesh_* m, int kind, const int* dims, int p, unsigned long long seed);
}

using uhm::Element_;
using uhm::Mesh_;
using uhm::Node_;

namespace {
thread_local std::string g_last_error;
int g_num_threads = 1;
int g_block = 64;
double g_flop = 0.0;

Mesh_* M(uhm_fort_p* h) {
  if (!h || !*h) throw std::runtime_error("null mesh handle");
  return reinterpret_cast<Mesh_*>(*h);
}
Element_* E(uhm_fort_p* h) {
  if (!h || !*h) throw std::runtime_error("null element handle");
  return reinterpret_cast<Element_*>(*h);
}
Node_* N(uhm_fort_p* h) {
  if (!h || !*h) throw std::runtime_error("null node handle");
  return reinterpret_cast<Node_*>(*h);
}
uhm::FrontSolver* S(Mesh_* m) {
  if (!m->solver()) throw std::runtime_error("fronts not created (create_matrix_without_buffer)");
  return m->solver();
}

template <class F>
void guard(const char* name, F&& f) {
  try {
    f();
  } catch (const std::exception& e) {
    g_last_error = std::string(name) + ": " + e.what();
  } catch (...) {
    g_last_error = std::string(name) + ": unknown error";
  }
}

void factor(Mesh_* m, int method) {
  uhm::FrontSolver* s = S(m);
  s->set_block(g_block);
  m->last_info = s->factor(method);
  double fd, fsv, nnz, buf;
  m->estimate_cost(method, uhm::UHM_REAL, s->n_rhs(), fd, fsv, nnz, buf);
  m->last_flop = fd;
  g_flop += fd;
}

void random(Mesh_* m, bool spd) {
  uhm::FrontSolver* s = S(m);
  const uhm::RandomParams& r = m->rnd;
  s->generate(r.seed, r.alpha, spd ? 0.0 : r.beta, r.scale_kind, r.p0, r.p1, 1, 1);
}
}  // namespace

extern "C" {

void uhm_initialize_fla_(void) {}
void uhm_finalize_fla_(void) {}
void uhm_set_num_threads_(uhm_fort_int* n) { g_num_threads = (n && *n > 0) ? *n : 1; }
void uhm_get_num_threads_(uhm_fort_int* n) { *n = g_num_threads; }
void uhm_set_hier_block_size_(uhm_fort_int* b) {
  guard("uhm_set_hier_block_size", [&] {
    if (*b < 1) throw std::runtime_error("block size < 1");
    g_block = *b > 64 ? 64 : *b;
  });
}
void uhm_get_hier_block_size_(uhm_fort_int* b) { *b = g_block; }
void uhm_reset_flop_(void) { g_flop = 0.0; }
void uhm_get_flop_(uhm_fort_double* f) { *f = g_flop; }
void uhm_get_buffer_used_(uhm_fort_double* b) { *b = (double)uhm::device_bytes_used(); }
void uhm_get_max_buffer_used_(uhm_fort_double* b) { *b = (double)uhm::device_bytes_max(); }
void uhm_is_multithreading_enable_(uhm_fort_int* f) { *f = 1; }
void uhm_is_hier_matrix_enable_(uhm_fort_int* f) { *f = 1; }
void uhm_timer_(uhm_fort_double* t) {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  *t = (double)tv.tv_sec + 1.0e-6 * (double)tv.tv_usec;
}

void uhm_mesh_create_(uhm_fort_p* mesh) {
  guard("uhm_mesh_create", [&] { *mesh = reinterpret_cast<uhm_fort_p>(new Mesh_()); });
}
void uhm_mesh_delete_(uhm_fort_p* mesh) {
  guard("uhm_mesh_delete", [&] {
    delete M(mesh);
    *mesh = 0;
  });
}
void uhm_add_element_(uhm_fort_p* mesh, uhm_fort_p* elt) {
  guard("uhm_add_element", [&] { *elt = reinterpret_cast<uhm_fort_p>(M(mesh)->add_element()); });
}
void uhm_add_node_(uhm_fort_p* mesh, uhm_fort_int* id, uhm_fort_int* phy, uhm_fort_int* n_dof, uhm_fort_int* p,
                   uhm_fort_p* nod) {
  guard("uhm_add_node", [&] {
    *nod = reinterpret_cast<uhm_fort_p>(M(mesh)->add_node(*id, *phy, *n_dof, *p, uhm::UHM_NODE_KIND_DEFAULT));
  });
}
void uhm_element_add_node_(uhm_fort_p* elt, uhm_fort_p* nod) {
  guard("uhm_element_add_node", [&] { E(elt)->add_node(N(nod)); });
}
void uhm_copy_in_(uhm_fort_p* mesh, uhm_fort_p* elt, uhm_fort_int* datatype, uhm_fort_int* mm, uhm_fort_int* nn,
                  uhm_fort_int* node_disp, uhm_fort_int* nods, uhm_fort_int* side, uhm_fort_double* buffer) {
  guard("uhm_copy_in",
        [&] { M(mesh)->copy_in(E(elt), *datatype, *mm, *nn, *node_disp, nods, *side, buffer); });
}
void uhm_copy_out_(uhm_fort_p* mesh, uhm_fort_p* elt, uhm_fort_int* datatype, uhm_fort_int* mm, uhm_fort_int* nn,
                   uhm_fort_int* node_disp, uhm_fort_int* nods, uhm_fort_int* side, uhm_fort_double* buffer) {
  guard("uhm_copy_out",
        [&] { M(mesh)->copy_out(E(elt), *datatype, *mm, *nn, *node_disp, nods, *side, buffer); });
}
void uhm_mesh_disp_(uhm_fort_p* mesh) {
  guard("uhm_mesh_disp", [&] {
    Mesh_* m = M(mesh);
    printf("- Mesh -\n  nodes %d, elements %d, locked %d, dofs %lld\n", m->get_n_nodes(), m->get_n_elements(),
           (int)m->is_locked(), m->get_n_dof());
  });
}
void uhm_get_n_dof_(uhm_fort_p* mesh, uhm_fort_int* n) {
  guard("uhm_get_n_dof", [&] { *n = (int)M(mesh)->get_n_dof(); });
}
void uhm_get_n_nonzero_(uhm_fort_p* mesh, uhm_fort_int* n) {
  guard("uhm_get_n_nonzero", [&] { *n = (int)M(mesh)->get_n_nonzero(); });
}
void uhm_estimate_cost_(uhm_fort_p* mesh, uhm_fort_int* method, uhm_fort_int* datatype, uhm_fort_int* n_rhs,
                        uhm_fort_double* fd, uhm_fort_double* fsv, uhm_fort_uint* nnz, uhm_fort_double* buffer) {
  guard("uhm_estimate_cost", [&] {
    double a, b, c, d;
    M(mesh)->estimate_cost(*method, *datatype, *n_rhs, a, b, c, d);
    *fd = a;
    *fsv = b;
    *nnz = (unsigned int)c;
    *buffer = d;
  });
}
void uhm_mesh_import_(uhm_fort_p* mesh, uhm_fort_char* filename) {
  guard("uhm_mesh_import", [&] { M(mesh)->import_file(filename); });
}
void uhm_build_tree_(uhm_fort_p* mesh) {
  guard("uhm_build_tree", [&] { M(mesh)->build_tree_coarsening(); });
}
void uhm_create_matrix_without_buffer_(uhm_fort_p* mesh, uhm_fort_int* datatype, uhm_fort_int* n_rhs) {
  guard("uhm_create_matrix_without_buffer", [&] { M(mesh)->create_matrix_without_buffer(*datatype, *n_rhs); });
}
void uhm_create_leaf_matrix_buffer_(uhm_fort_p* mesh) {
  guard("uhm_create_leaf_matrix_buffer", [&] { M(mesh)->create_matrix_buffer(1); });
}
void uhm_create_element_matrix_buffer_(uhm_fort_p* mesh, uhm_fort_int* is_schur) {
  guard("uhm_create_element_matrix_buffer", [&] { M(mesh)->create_matrix_buffer(*is_schur); });
}
void uhm_set_rhs_(uhm_fort_p* mesh) {
  guard("uhm_set_rhs", [&] { M(mesh)->set_rhs(); });
}
void uhm_random_spd_matrix_(uhm_fort_p* mesh) {
  guard("uhm_random_spd_matrix", [&] { random(M(mesh), true); });
}
void uhm_random_matrix_(uhm_fort_p* mesh) {
  guard("uhm_random_matrix", [&] { random(M(mesh), false); });
}
void uhm_triangularize_(uhm_fort_p* mesh) {
  guard("uhm_triangularize", [&] { (void)M(mesh); });
}
void uhm_set_node_(uhm_fort_p* mesh, uhm_fort_int* id, uhm_fort_int* phy, uhm_fort_int* n_dof, uhm_fort_int* p) {
  guard("uhm_set_node", [&] { M(mesh)->set_node(*id, *phy, *n_dof, *p); });
}
void uhm_backup_(uhm_fort_p* mesh, uhm_fort_p* backup) {
  guard("uhm_backup", [&] { M(mesh)->backup(*M(backup)); });
}
void uhm_check_reuse_(uhm_fort_p* mesh, uhm_fort_p* backup) {
  guard("uhm_check_reuse", [&] { M(mesh)->check_reuse(*M(backup)); });
}
void uhm_get_n_reusable_(uhm_fort_p* mesh, uhm_fort_int* n) {
  guard("uhm_get_n_reusable", [&] { *n = M(mesh)->n_reusable(); });
}
void uhm_get_n_reused_(uhm_fort_p* mesh, uhm_fort_int* n) {
  guard("uhm_get_n_reused", [&] { *n = S(M(mesh))->n_reused(); });
}
void uhm_lock_(uhm_fort_p* mesh) {
  guard("uhm_lock", [&] { M(mesh)->lock(); });
}
void uhm_unlock_(uhm_fort_p* mesh) {
  guard("uhm_unlock", [&] { M(mesh)->unlock(); });
}
void uhm_chol_with_free_(uhm_fort_p* mesh) {
  guard("uhm_chol_with_free", [&] { factor(M(mesh), uhm::UHM_CHOL); });
}
void uhm_lu_piv_with_free_(uhm_fort_p* mesh) {
  guard("uhm_lu_piv_with_free", [&] { factor(M(mesh), uhm::UHM_LU_PIV); });
}
void uhm_solve_chol_(uhm_fort_p* mesh) {
  guard("uhm_solve_chol", [&] { S(M(mesh))->solve(uhm::UHM_CHOL); });
}
void uhm_solve_lu_piv_(uhm_fort_p* mesh) {
  guard("uhm_solve_lu_piv", [&] { S(M(mesh))->solve(uhm::UHM_LU_PIV); });
}
void uhm_check_chol_(uhm_fort_p* mesh) {
  guard("uhm_check_chol", [&] { S(M(mesh))->check(uhm::UHM_CHOL); });
}
void uhm_check_lu_piv_(uhm_fort_p* mesh) {
  guard("uhm_check_lu_piv", [&] { S(M(mesh))->check(uhm::UHM_LU_PIV); });
}
void uhm_get_residual_(uhm_fort_p* mesh, uhm_fort_double* res) {
  guard("uhm_get_residual", [&] { *res = S(M(mesh))->get_residual(); });
}

// ---- extensions
void uhm_element_set_parent_(uhm_fort_p* elt, uhm_fort_p* parent) {
  guard("uhm_element_set_parent", [&] { E(elt)->set_parent(E(parent)); });
}
void uhm_element_add_child_(uhm_fort_p* elt, uhm_fort_p* child) {
  guard("uhm_element_add_child", [&] { E(elt)->add_child(E(child)); });
}
void uhm_build_synthetic_mesh_(uhm_fort_p* mesh, uhm_fort_int* kind, uhm_fort_int* dims, uhm_fort_int* p,
                               long long* seed) {
  guard("uhm_build_synthetic_mesh",
        [&] { uhm::build_synthetic_mesh(M(mesh), *kind, dims, *p, (unsigned long long)*seed); });
}
void uhm_set_random_params_(uhm_fort_p* mesh, long long* seed, uhm_fort_double* alpha, uhm_fort_double* beta,
                            uhm_fort_int* scale_kind, uhm_fort_double* p0, uhm_fort_double* p1) {
  guard("uhm_set_random_params", [&] {
    uhm::RandomParams& r = M(mesh)->rnd;
    r.seed = (unsigned long long)*seed;
    r.alpha = *alpha;
    r.beta = *beta;
    r.scale_kind = *scale_kind;
    r.p0 = *p0;
    r.p1 = *p1;
  });
}
void uhm_get_info_(uhm_fort_p* mesh, uhm_fort_int* info) {
  guard("uhm_get_info", [&] { *info = M(mesh)->last_info; });
}
void uhm_get_backward_error_(uhm_fort_p* mesh, uhm_fort_double* eta) {
  guard("uhm_get_backward_error", [&] { *eta = S(M(mesh))->backward_error(); });
}
void uhm_get_timing_(uhm_fort_p* mesh, uhm_fort_double* fms, uhm_fort_double* sms) {
  guard("uhm_get_timing", [&] {
    *fms = S(M(mesh))->last_factor_ms();
    *sms = S(M(mesh))->last_solve_ms();
  });
}
void uhm_set_precision_(uhm_fort_p* mesh, uhm_fort_int* mode, uhm_fort_int* top_levels) {
  guard("uhm_set_precision", [&] { S(M(mesh))->set_precision(*mode, *top_levels); });
}
void uhm_improve_chol_(uhm_fort_p* mesh, uhm_fort_double* tol, uhm_fort_int* max_iter, uhm_fort_int* n_iter) {
  guard("uhm_improve_chol", [&] { *n_iter = S(M(mesh))->refine(uhm::UHM_CHOL, *tol, *max_iter); });
}
void uhm_improve_lu_piv_(uhm_fort_p* mesh, uhm_fort_double* tol, uhm_fort_int* max_iter, uhm_fort_int* n_iter) {
  guard("uhm_improve_lu_piv", [&] { *n_iter = S(M(mesh))->refine(uhm::UHM_LU_PIV, *tol, *max_iter); });
}
void uhm_get_refine_history_(uhm_fort_p* mesh, uhm_fort_double* eta, uhm_fort_int* n, uhm_fort_double* ms) {
  guard("uhm_get_refine_history", [&] {
    const std::vector<double>& h = S(M(mesh))->refine_history();
    const int cap = *n;
    for (int k = 0; k < cap && k < (int)h.size(); ++k) eta[k] = h[k];
    *n = (int)h.size();
    *ms = S(M(mesh))->last_refine_ms();
  });
}
void uhm_set_refinement_(uhm_fort_p* mesh, uhm_fort_int* kind, uhm_fort_int* restart, uhm_fort_double* inner_tol) {
  guard("uhm_set_refinement", [&] { S(M(mesh))->set_refinement(*kind, *restart, *inner_tol); });
}
void uhm_get_refine_solves_(uhm_fort_p* mesh, uhm_fort_int* n) {
  guard("uhm_get_refine_solves", [&] { *n = S(M(mesh))->refine_solves(); });
}
void uhm_set_tuning_(uhm_fort_p* mesh, uhm_fort_int* key, uhm_fort_int* value) {
  guard("uhm_set_tuning", [&] { S(M(mesh))->set_tuning(*key, *value); });
}
void uhm_get_tuning_(uhm_fort_p* mesh, uhm_fort_int* key, uhm_fort_int* value) {
  guard("uhm_get_tuning", [&] { *value = S(M(mesh))->get_tuning(*key); });
}
void uhm_autotune_(uhm_fort_p* mesh, uhm_fort_int* method, uhm_fort_double* best_ms) {
  guard("uhm_autotune", [&] { *best_ms = S(M(mesh))->autotune(*method); });
}
void uhm_set_graph_(uhm_fort_p* mesh, uhm_fort_int* on) {
  guard("uhm_set_graph", [&] { S(M(mesh))->set_graph(*on != 0); });
}
void uhm_set_profile_(uhm_fort_p* mesh, uhm_fort_int* on) {
  guard("uhm_set_profile", [&] { S(M(mesh))->set_profile(*on != 0); });
}
void uhm_get_kernel_stats_(uhm_fort_p* mesh, uhm_fort_double* ms, uhm_fort_double* flops, uhm_fort_double* bytes,
                           uhm_fort_int* launches) {
  guard("uhm_get_kernel_stats", [&] {
    const uhm::KernelStats& s = S(M(mesh))->kernel_stats();
    for (int k = 0; k < uhm::N_LAUNCH_KINDS; ++k) {
      ms[k] = s.ms[k];
      flops[k] = s.flops[k];
      bytes[k] = s.bytes[k];
      launches[k] = s.launches[k];
    }
  });
}
int uhm_get_last_error_(char* buf, int len) {
  const int n = (int)g_last_error.size();
  if (buf && len > 0) {
    const int c = n < len - 1 ? n : len - 1;
    memcpy(buf, g_last_error.data(), c);
    buf[c] = 0;
  }
  g_last_error.clear();
  return n;
}
void uhm_get_n_fronts_(uhm_fort_p* mesh, uhm_fort_int* n) {
  guard("uhm_get_n_fronts", [&] { *n = S(M(mesh))->n_fronts(); });
}
void uhm_get_front_info_(uhm_fort_p* mesh, uhm_fort_int* front, uhm_fort_int* out) {
  guard("uhm_get_front_info", [&] {
    uhm::FrontSolver* s = S(M(mesh));
    if (*front < 0 || *front >= s->n_fronts()) throw std::runtime_error("front index out of range");
    Element_* e = s->element_of(*front);
    out[0] = e->id;
    out[1] = e->fs;
    out[2] = e->ss;
    out[3] = s->level_of(*front);
    out[4] = s->front(*front).parent;
    out[5] = e->generation;
    out[6] = (int)e->factor.size();
    out[7] = (int)e->schur.size();
    out[8] = e->is_leaf() ? 1 : 0;
  });
}
void uhm_get_front_nodes_(uhm_fort_p* mesh, uhm_fort_int* front, uhm_fort_int* which, uhm_fort_int* ids,
                          uhm_fort_int* offs, uhm_fort_int* n) {
  guard("uhm_get_front_nodes", [&] {
    uhm::FrontSolver* s = S(M(mesh));
    if (*front < 0 || *front >= s->n_fronts()) throw std::runtime_error("front index out of range");
    Element_* e = s->element_of(*front);
    const auto& L = (*which == 0) ? e->factor : e->schur;
    *n = (int)L.size();
    if (ids)
      for (size_t i = 0; i < L.size(); ++i) {
        ids[i] = L[i].first->id;
        if (offs) offs[i] = L[i].second;
      }
  });
}
void uhm_get_mapper_(uhm_fort_p* mesh, uhm_fort_int* child_front, uhm_fort_int* runs, uhm_fort_int* n) {
  guard("uhm_get_mapper", [&] {
    uhm::FrontSolver* s = S(M(mesh));
    if (*child_front < 0 || *child_front >= s->n_fronts()) throw std::runtime_error("front index out of range");
    Element_* c = s->element_of(*child_front);
    if (!c->parent) {
      *n = 0;
      return;
    }
    std::vector<uhm::Mapper_> r = uhm::build_mapper(c->parent, c);
    *n = (int)r.size();
    if (runs)
      for (size_t i = 0; i < r.size(); ++i) {
        runs[4 * i + 0] = r[i].offs_c;
        runs[4 * i + 1] = r[i].offs_p;
        runs[4 * i + 2] = r[i].fs_p;
        runs[4 * i + 3] = r[i].n_dof;
      }
  });
}
void uhm_get_node_offset_(uhm_fort_p* mesh, uhm_fort_int* id, uhm_fort_int* phy, uhm_fort_int* offset) {
  guard("uhm_get_node_offset", [&] {
    Node_* nd = M(mesh)->find_node(*id, *phy);
    if (!nd) throw std::runtime_error("unknown node");
    *offset = nd->offset;
  });
}
void uhm_fetch_front_(uhm_fort_p* mesh, uhm_fort_int* front, uhm_fort_double* out) {
  guard("uhm_fetch_front", [&] { S(M(mesh))->fetch_front_raw(*front, out); });
}
void uhm_fetch_pivots_(uhm_fort_p* mesh, uhm_fort_int* front, uhm_fort_int* out) {
  guard("uhm_fetch_pivots", [&] { S(M(mesh))->fetch_pivots(*front, out); });
}
void uhm_fetch_stage_(uhm_fort_p* mesh, uhm_fort_int* front, uhm_fort_double* out) {
  guard("uhm_fetch_stage", [&] { S(M(mesh))->fetch_stage(*front, out); });
}
void uhm_fetch_rhs_front_(uhm_fort_p* mesh, uhm_fort_int* front, uhm_fort_double* out) {
  guard("uhm_fetch_rhs_front", [&] { S(M(mesh))->fetch_rhs_front(*front, out); });
}
void uhm_fetch_x_global_(uhm_fort_p* mesh, uhm_fort_double* out) {
  guard("uhm_fetch_x_global", [&] { S(M(mesh))->fetch_x_global(out); });
}
void uhm_fetch_b_global_(uhm_fort_p* mesh, uhm_fort_double* out) {
  guard("uhm_fetch_b_global", [&] { S(M(mesh))->fetch_b_global(out); });
}
void uhm_export_connectivity_(uhm_fort_p* mesh, uhm_fort_int* sizes, uhm_fort_int* node_id, uhm_fort_int* ndof,
                              uhm_fort_int* kind, uhm_fort_int* eptr, uhm_fort_int* enode, uhm_fort_int* parent,
                              uhm_fort_int* cptr, uhm_fort_int* cidx) {
  guard("uhm_export_connectivity", [&] {
    Mesh_* m = M(mesh);
    const int nn = m->get_n_nodes(), ne = m->get_n_elements();
    long long ln = 0, lc = 0;
    for (int e = 0; e < ne; ++e) {
      Element_* el = m->element_at(e);
      if (el->is_leaf()) ln += (long long)el->nodes.size();
      lc += (long long)el->children.size();
    }
    sizes[0] = nn;
    sizes[1] = ne;
    sizes[2] = (int)ln;
    sizes[3] = (int)lc;
    if (!node_id) return;
    for (int i = 0; i < nn; ++i) {
      Node_* nd = m->node_at(i);
      node_id[i] = nd->id;
      ndof[i] = nd->n_dof;
      kind[i] = nd->kind;
    }
    int pn = 0, pc = 0;
    eptr[0] = 0;
    cptr[0] = 0;
    for (int e = 0; e < ne; ++e) {
      Element_* el = m->element_at(e);
      if (el->is_leaf())
        for (auto& nf : el->nodes) enode[pn++] = nf.first->seq;
      eptr[e + 1] = pn;
      parent[e] = el->parent ? el->parent->id : -1;
      for (Element_* c : el->children) cidx[pc++] = c->id;
      cptr[e + 1] = pc;
    }
  });
}
void uhm_get_csr_size_(uhm_fort_p* mesh, long long* nnz) {
  guard("uhm_get_csr_size", [&] {
    std::vector<long long> rp;
    std::vector<int> c;
    std::vector<double> v;
    S(M(mesh))->csr(rp, c, v);
    *nnz = (long long)c.size();
  });
}
void uhm_fetch_csr_(uhm_fort_p* mesh, long long* rowptr, uhm_fort_int* col, uhm_fort_double* val) {
  guard("uhm_fetch_csr", [&] {
    std::vector<long long> rp;
    std::vector<int> c;
    std::vector<double> v;
    S(M(mesh))->csr(rp, c, v);
    std::copy(rp.begin(), rp.end(), rowptr);
    std::copy(c.begin(), c.end(), col);
    std::copy(v.begin(), v.end(), val);
  });
}

}  // extern "C"
