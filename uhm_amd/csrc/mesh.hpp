// Host-side mesh, element tree and symbolic analysis (lock) of the
// multifrontal solver.  This is the drop-in C++ API surface: a caller
// builds nodes and elements, builds (or supplies) the element tree, locks
// the mesh, creates the fronts, copies element matrices in, factorizes,
// solves and checks.  Numeric work is delegated to FrontSolver, which runs
// every front operation on the GPU.
//
// Reference interfaces mirrored here:
//   Mesh_     UHM/include/uhm/mesh/mesh.hxx:58-301
//   Element_  UHM/include/uhm/mesh/element.hxx:48-163
//   Node_     UHM/include/uhm/mesh/node.hxx:45-102
//   Helper_/Mapper_ (extend-add index maps) UHM/include/uhm/matrix/uhm/helper.hxx:43-158
#pragma once

#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace uhm {

// Constants with the reference's numeric values (UHM/include/uhm/const.hxx:70-81,
// libflame datatype codes behind LINAL_REAL / LINAL_SINGLE_REAL).
enum { UHM_PHYSICS_SINGLE = 1, UHM_PHYSICS_MULTI };
enum { UHM_NODE_KIND_DEFAULT = 0, UHM_NODE_KIND_BOUNDARY };
enum { UHM_LHS = 1, UHM_RHS };
enum { UHM_CHOL = 1, UHM_LU_NOPIV, UHM_LU_PIV, UHM_LU_INCPIV, UHM_QR };
enum { UHM_NOT_SEPARATED = 0, UHM_SEPARATED_FACTOR, UHM_SEPARATED_SCHUR };
constexpr int UHM_SINGLE_REAL = 101;  // FLA_FLOAT
constexpr int UHM_REAL = 102;         // FLA_DOUBLE (LINAL_REAL)

class Element_;
class Mesh_;
class FrontSolver;

// One run of the extend-add map of a (parent, child) pair, with the
// reference's field meaning (helper.hxx:43): child schur offset, parent
// offset inside the parent's factor (fs_p = 0) or schur (fs_p = 1) list,
// and the number of consecutive dofs.
struct Mapper_ {
  int offs_c, offs_p, fs_p, n_dof;
};

class Node_ {
 public:
  int id = 0, phy = 0;
  int n_dof = 0, p = 0, kind = 0;
  int offset = -1;  // global offset assigned by numbering; -1 = not numbered
  int seq = 0;      // creation index inside the mesh
  int marker = 0;
  std::vector<Element_*> owner;  // set semantics (node.hxx:47)

  int get_n_owner() const { return (int)owner.size(); }
  void add_owner(Element_* e);
  void remove_owner(Element_* e);
  void clean_connectivity();  // drop non-leaf owners (node.cxx:98-106)
};

class Element_ {
 public:
  int id = 0;
  int generation = 0;
  Element_* parent = nullptr;
  std::vector<Element_*> children;
  // (node, separation flag), kept sorted by node creation index.  The
  // reference keys this map by Node_* address (element.hxx:58), which makes
  // the element-local order allocator dependent; creation order is the
  // deterministic stand-in.
  std::vector<std::pair<Node_*, int>> nodes;
  // (node, local offset), sorted by global offset (element.cxx:288-334).
  std::vector<std::pair<Node_*, int>> factor, schur;
  int fs = 0, ss = 0;       // factor / schur dof counts after arrange
  int front = -1;           // index of this element's front in the solver
  bool reuse = false;

  bool is_leaf() const { return children.empty(); }
  bool is_orphan() const { return parent == nullptr; }
  void add_node(Node_* n, int separated = UHM_NOT_SEPARATED);
  int find_node(const Node_* n) const;  // index into nodes or -1
  void add_child(Element_* c) { children.push_back(c); }
  void set_parent(Element_* p) { parent = p; }
  void update_generation();
  void separate_nodes();
  void merge_nodes_from_children();
  void restore_connectivity();
  void numbering(long long& g_offset);
  void arrange_nodes();
  std::pair<int, int> get_n_dof() const;
  void estimate_cost(int method, int datatype, int n_rhs, double& flop_decompose,
                     double& flop_solve, double& n_nonzero_factor, double& buffer) const;
};

// Build the (parent, child) map like Helper_::set_mapper (helper.hxx:98-158):
// per child schur node, its place in the parent's factor or schur list,
// condensed into runs of contiguous parent offsets.
std::vector<Mapper_> build_mapper(const Element_* parent, const Element_* child);

// Parameters of the device element-matrix generator (random_matrix).
struct RandomParams {
  unsigned long long seed = 1;
  double alpha = 0.0, beta = 0.0;
  int scale_kind = 0;
  double p0 = 0.0, p1 = 0.0;
};

class Mesh_ {
 public:
  RandomParams rnd;
  int last_info = 0;       // status of the last factorization
  double last_flop = 0.0;  // algorithmic flops of the last factorization

  Mesh_();
  ~Mesh_();

  // --- mesh construction (mesh.cxx:268-349)
  Node_* add_node(int id, int phy, int n_dof, int p, int kind);
  Node_* find_node(int id, int phy) const;
  // new dof count and order of an existing node (Node_::set_n_dof / set_p,
  // node.hxx:72-73), mesh unlocked
  void set_node(int id, int phy, int n_dof, int p);
  Element_* add_element(int gen = 0);
  Element_* find_element(int id) const;
  int get_n_nodes() const { return (int)nodes_.size(); }
  int get_n_elements() const { return (int)elements_.size(); }
  Element_* get_root() const;
  Element_* element_at(int i) { return &elements_[i]; }
  Node_* node_at(int i) { return &nodes_[i]; }
  bool import_file(const char* path);  // ".uhm" text format (mesh.cxx:511-579)
  void remove_orphan_nodes();          // nodes no element holds (mesh.cxx:246-251)

  // --- tree + analysis
  void update_generation();  // leaf 0, parent = min(children) - 1 (element.cxx:63-72)
  void build_tree_coarsening();  // METIS-free tree builder (build_tree_var1.cxx:53-181)
  void build_tree_bisection();   // breadth-first bisection of the orphans (disconnected meshes)
  void lock();    // mesh.cxx:438-456
  void unlock();
  bool is_locked() const { return locked_; }

  // --- scheduler view (scheduler.cxx:157-178): generations deepest first
  const std::vector<std::vector<Element_*>>& levels() const { return levels_; }
  const std::vector<Element_*>& leaves() const { return leaves_; }

  // --- fronts
  void create_matrix_without_buffer(int datatype, int n_rhs);
  void create_matrix_buffer(int is_schur);
  void free_matrix();
  bool is_matrix_created() const { return solver_ != nullptr; }
  FrontSolver* solver() { return solver_.get(); }

  // --- element matrices / right-hand sides (interf.cxx:108-204)
  void copy_in(Element_* e, int datatype, int m, int n, int nod_disp, const int* nods,
               int side, const double* buffer);
  void copy_out(Element_* e, int datatype, int m, int n, int nod_disp, const int* nods,
                int side, double* buffer);
  void set_rhs();

  // --- factorization reuse (mesh.cxx:113-225)
  // b keeps this mesh's connectivity: per element id its factor and schur
  // node lists as (node, n_dof + p + kind), as Mesh_::backup stores them
  // (b's own elements, nodes and fronts are dropped first, as by reset()).
  void backup(Mesh_& b) const;
  // Reuse flag of every element against the connectivity b keeps: set where
  // the element exists there with the same factor and schur lists (same
  // nodes, same encoded values), cleared on every ancestor of an element
  // without it.  The next create_matrix_without_buffer keeps the factored
  // fronts of the reusable elements for the next factorization.
  void check_reuse(const Mesh_& b);
  int n_reusable() const;
  void reset();

  // --- global quantities
  long long get_n_dof() const;           // mesh/matrix.cxx:189-200
  long long get_n_nonzero() const;       // mesh/matrix.cxx:213-226
  long long get_n_nonzero_factor() const;
  void estimate_cost(int method, int datatype, int n_rhs, double& flop_decompose,
                     double& flop_solve, double& n_nonzero_factor, double& buffer) const;

 private:
  std::deque<Node_> nodes_;
  std::unordered_map<long long, Node_*> node_index_;
  std::deque<Element_> elements_;
  int id_element_ = 0;
  bool locked_ = false;
  std::vector<std::vector<Element_*>> levels_;  // index 0 = generation 0 (leaves)
  std::vector<int> level_gen_;
  std::vector<Element_*> leaves_;
  std::unique_ptr<FrontSolver> solver_;
  struct ReuseRecord {
    std::vector<std::pair<const Node_*, int>> factor, schur;
  };
  std::unordered_map<int, ReuseRecord> backup_;  // element id -> connectivity kept by backup()
  friend class FrontSolver;
};

// LAWN-41 flop counts with the reference's formulas (LINAL/src/util.cxx:56-101).
double flop_gemm(double m, double n, double k);
double flop_syrk(double k, double n);
double flop_trsm_lower(double m, double n);
double flop_trsm_upper(double m, double n);
double flop_lu(double m, double n);
double flop_chol(double n);

}  // namespace uhm
