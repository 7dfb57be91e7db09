/*
 * C++ interface of the multifrontal solver for programs written against the
 * reference's uhm::Mesh_ (UHM/include/uhm/mesh/mesh.hxx:52-300,
 * element.hxx:44-140, node.hxx:40-100): the same class and method names,
 * argument meanings and call sequence (create nodes and elements, build the
 * tree, lock, create the fronts, copy in, factorize, solve, check, copy
 * out), implemented inline over the C ABI of uhm_amd.h, so such a program
 * recompiles against this header and links against libuhm_amd.so.
 *
 * Handles follow the reference's pointer typedefs (Mesh = Mesh_*, Element =
 * Element_*, Node = Node_*); nodes and elements are owned by their mesh.
 * Errors raise std::runtime_error with the library's message (where the
 * reference asserts).  Operations the GPU path does not provide (QR, LU
 * without pivoting, out-of-core variants, the split 1 / 2 solve sweeps,
 * export to external sparse solvers) are not declared.
 */
#ifndef UHM_AMD_MESH_HPP_
#define UHM_AMD_MESH_HPP_

#include <deque>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "uhm_amd.h"

namespace uhm {

constexpr int UHM_REAL = UHM_AMD_REAL;
constexpr int UHM_CHOL = UHM_AMD_CHOL;
constexpr int UHM_LU_PIV = UHM_AMD_LU_PIV;
constexpr int UHM_LHS = UHM_AMD_LHS;
constexpr int UHM_RHS = UHM_AMD_RHS;
constexpr int UHM_PHYSICS_SINGLE = UHM_AMD_PHYSICS_SINGLE;
constexpr int UHM_PHYSICS_MULTI = UHM_AMD_PHYSICS_MULTI;

namespace detail {
inline void check() {
  char buf[1024];
  if (uhm_get_last_error_(buf, (int)sizeof(buf)) > 0) throw std::runtime_error(buf);
}
}  // namespace detail

class Mesh_;

// node.hxx:40-100: id (id, physics), dofs, polynomial order
class Node_ {
 public:
  std::pair<int, int> get_id() const { return id_; }
  int get_n_dof() const { return n_dof_; }
  int get_p() const { return p_; }

 private:
  friend class Mesh_;
  friend class Element_;
  uhm_fort_p h_ = 0;
  std::pair<int, int> id_{0, 0};
  int n_dof_ = 0, p_ = 0;
};
typedef Node_* Node;

// element.hxx:44-140: nodes, parent and children of the element tree
class Element_ {
 public:
  void add_node(Node n) {
    uhm_element_add_node_(&h_, &n->h_);
    detail::check();
  }
  void set_parent(Element_* p) {
    uhm_element_set_parent_(&h_, &p->h_);
    detail::check();
  }
  void add_child(Element_* c) {
    uhm_element_add_child_(&h_, &c->h_);
    detail::check();
  }
  int get_id() const { return id_; }

 private:
  friend class Mesh_;
  uhm_fort_p h_ = 0;
  int id_ = 0;
};
typedef Element_* Element;

class Mesh_ {
 public:
  Mesh_() {
    uhm_mesh_create_(&h_);
    detail::check();
  }
  ~Mesh_() { uhm_mesh_delete_(&h_); }
  Mesh_(const Mesh_&) = delete;
  Mesh_& operator=(const Mesh_&) = delete;

  // --- construction (mesh.cxx:268-349)
  Node add_node(int id, int n_dof) { return add_node(std::make_pair(id, 0), n_dof, n_dof); }
  Node add_node(int id, int n_dof, int p) { return add_node(std::make_pair(id, 0), n_dof, p); }
  Node add_node(std::pair<int, int> id, int n_dof) { return add_node(id, n_dof, n_dof); }
  Node add_node(std::pair<int, int> id, int n_dof, int p) {
    int i = id.first, phy = id.second;
    uhm_fort_p h = 0;
    uhm_add_node_(&h_, &i, &phy, &n_dof, &p, &h);
    detail::check();
    nodes_.emplace_back();
    Node_& n = nodes_.back();
    n.h_ = h;
    n.id_ = id;
    n.n_dof_ = n_dof;
    n.p_ = p;
    return &n;
  }
  Element add_element() {
    uhm_fort_p h = 0;
    uhm_add_element_(&h_, &h);
    detail::check();
    elements_.emplace_back();
    Element_& e = elements_.back();
    e.h_ = h;
    e.id_ = (int)elements_.size() - 1;
    return &e;
  }
  int get_n_elements() const { return (int)elements_.size(); }
  bool import_file(const char* full_path) {
    uhm_mesh_import_(&h_, const_cast<char*>(full_path));
    detail::check();
    nodes_.clear();
    elements_.clear();
    return true;
  }

  // --- factorization reuse (mesh.cxx:113-225)
  void backup(Mesh_* b) {
    uhm_backup_(&h_, &b->h_);
    detail::check();
  }
  void check_reuse(Mesh_* b) {
    uhm_check_reuse_(&h_, &b->h_);
    detail::check();
  }

  // --- solving sequence (mesh.hxx:170-300)
  void lock() { call(uhm_lock_); }
  void unlock() { call(uhm_unlock_); }
  void create_matrix_without_buffer(int datatype, int n_rhs) {
    uhm_create_matrix_without_buffer_(&h_, &datatype, &n_rhs);
    detail::check();
  }
  void create_leaf_matrix_buffer() { call(uhm_create_leaf_matrix_buffer_); }
  void create_element_matrix_buffer() { create_element_matrix_buffer(0); }
  void create_element_matrix_buffer(int is_schur) {
    uhm_create_element_matrix_buffer_(&h_, &is_schur);
    detail::check();
  }
  void create_matrix_buffer() { create_element_matrix_buffer(0); }
  void create_matrix_buffer(int is_schur) { create_element_matrix_buffer(is_schur); }
  void random_matrix() { call(uhm_random_matrix_); }
  void random_spd_matrix() { call(uhm_random_spd_matrix_); }
  void triangularize() { call(uhm_triangularize_); }

  void copy_in(Element elt, int datatype, int m, int n, int* nods, int side, void* buffer) {
    copy_in(elt, datatype, m, n, UHM_PHYSICS_SINGLE, nods, side, buffer);
  }
  void copy_in(Element elt, int datatype, int m, int n, int nod_disp, int* nods, int side, void* buffer) {
    uhm_copy_in_(&h_, &elt->h_, &datatype, &m, &n, &nod_disp, nods, &side, static_cast<double*>(buffer));
    detail::check();
  }
  void copy_out(Element elt, int datatype, int m, int n, int* nods, int side, void* buffer) {
    copy_out(elt, datatype, m, n, UHM_PHYSICS_SINGLE, nods, side, buffer);
  }
  void copy_out(Element elt, int datatype, int m, int n, int nod_disp, int* nods, int side, void* buffer) {
    uhm_copy_out_(&h_, &elt->h_, &datatype, &m, &n, &nod_disp, nods, &side, static_cast<double*>(buffer));
    detail::check();
  }

  void set_rhs() { call(uhm_set_rhs_); }
  double get_residual() {
    double r = 0;
    uhm_get_residual_(&h_, &r);
    detail::check();
    return r;
  }
  unsigned int get_n_dof() {
    int n = 0;
    uhm_get_n_dof_(&h_, &n);
    detail::check();
    return (unsigned int)n;
  }
  unsigned int get_n_nonzero() {
    int n = 0;
    uhm_get_n_nonzero_(&h_, &n);
    detail::check();
    return (unsigned int)n;
  }
  void estimate_cost(int method, int datatype, int n_rhs, double& flop_decompose, double& flop_solve,
                     unsigned int& n_nonzero_factor, double& buffer) {
    uhm_estimate_cost_(&h_, &method, &datatype, &n_rhs, &flop_decompose, &flop_solve, &n_nonzero_factor, &buffer);
    detail::check();
  }

  // factorization, solve, check (the fronts stay in device memory either
  // way, so the _with_free / _without_free pairs run the same factorization)
  void chol_with_free() { call(uhm_chol_with_free_); }
  void chol_without_free() { call(uhm_chol_with_free_); }
  void solve_chol() { call(uhm_solve_chol_); }
  void check_chol() { call(uhm_check_chol_); }
  void lu_piv_with_free() { call(uhm_lu_piv_with_free_); }
  void lu_piv_without_free() { call(uhm_lu_piv_with_free_); }
  void solve_lu_piv() { call(uhm_solve_lu_piv_); }
  void check_lu_piv() { call(uhm_check_lu_piv_); }
  // iterative refinement to backward error tol (the reference's improve_*
  // are stubs); returns the corrections applied
  int improve_chol(double tol = 1e-12, int max_iter = 20) { return improve(uhm_improve_chol_, tol, max_iter); }
  int improve_lu_piv(double tol = 1e-12, int max_iter = 20) { return improve(uhm_improve_lu_piv_, tol, max_iter); }

  // the C ABI handle, for the extensions of uhm_amd.h
  uhm_fort_p* handle() { return &h_; }

 private:
  friend void build_tree(Mesh_* m);
  uhm_fort_p h_ = 0;
  std::deque<Node_> nodes_;
  std::deque<Element_> elements_;

  void call(void (*f)(uhm_fort_p*)) {
    f(&h_);
    detail::check();
  }
  int improve(void (*f)(uhm_fort_p*, uhm_fort_double*, uhm_fort_int*, uhm_fort_int*), double tol, int max_iter) {
    int n = 0;
    f(&h_, &tol, &max_iter, &n);
    detail::check();
    return n;
  }
};
typedef Mesh_* Mesh;

// The element tree over the mesh's elements (build_tree,
// UHM/src/mesh/build_tree_var1.cxx:53-181; here the METIS-free bisection).
inline void build_tree(Mesh m) {
  uhm_build_tree_(&m->h_);
  detail::check();
}

inline void initialize() { uhm_initialize_fla_(); }
inline void finalize() { uhm_finalize_fla_(); }

}  // namespace uhm

#endif
