// Host-side mesh API and symbolic analysis.  See mesh.hpp for the reference
// interfaces this mirrors; each routine below cites the reference code whose
// behaviour it restates.
#include "mesh.hpp"

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <set>
#include <stdexcept>

#include "solver.hpp"

namespace uhm {

static inline long long node_key(int id, int phy) {
  return ((long long)id << 32) ^ (unsigned int)phy;
}

// ---------------------------------------------------------------- Node_
void Node_::add_owner(Element_* e) {
  for (Element_* o : owner)
    if (o == e) return;
  owner.push_back(e);
}

void Node_::remove_owner(Element_* e) {
  for (size_t i = 0; i < owner.size(); ++i)
    if (owner[i] == e) {
      owner.erase(owner.begin() + i);
      return;
    }
}

void Node_::clean_connectivity() {
  // node.cxx:98-106: owners that are not leaves are dropped.
  std::vector<Element_*> keep;
  for (Element_* o : owner)
    if (o->is_leaf()) keep.push_back(o);
  owner.swap(keep);
}

// ---------------------------------------------------------------- Element_
int Element_::find_node(const Node_* n) const {
  auto it = std::lower_bound(nodes.begin(), nodes.end(), n->seq,
                             [](const std::pair<Node_*, int>& a, int s) { return a.first->seq < s; });
  if (it != nodes.end() && it->first == n) return (int)(it - nodes.begin());
  return -1;
}

void Element_::add_node(Node_* n, int separated) {
  // element.cxx:251-262: insert or overwrite the flag; a new member
  // registers this element as an owner of the node.
  auto it = std::lower_bound(nodes.begin(), nodes.end(), n->seq,
                             [](const std::pair<Node_*, int>& a, int s) { return a.first->seq < s; });
  if (it != nodes.end() && it->first == n) {
    it->second = separated;
    return;
  }
  nodes.insert(it, std::make_pair(n, separated));
  n->add_owner(this);
}

void Element_::update_generation() {
  // element.cxx:63-72
  if (is_leaf()) return;
  int gen = 0;
  for (Element_* c : children) gen = std::min(gen, c->generation);
  generation = gen - 1;
}

void Element_::separate_nodes() {
  // element.cxx:272-286: a node is eliminated here iff this element is its
  // only owner and it is not a boundary-kind node.
  for (auto& nf : nodes) {
    if (nf.first->get_n_owner() == 0) throw std::runtime_error("separate_nodes: node without owner");
    nf.second = (nf.first->get_n_owner() == 1 && nf.first->kind == UHM_NODE_KIND_DEFAULT)
                    ? UHM_SEPARATED_FACTOR
                    : UHM_SEPARATED_SCHUR;
  }
}

void Element_::merge_nodes_from_children() {
  // element.cxx:336-360: the parent's node set is the union of its children's
  // schur nodes; each child stops owning the nodes it passes up.
  if (is_leaf()) return;
  nodes.clear();
  std::vector<Node_*> cand;
  for (Element_* c : children)
    for (auto& nf : c->nodes)
      if (nf.second == UHM_SEPARATED_SCHUR) cand.push_back(nf.first);
  std::sort(cand.begin(), cand.end(), [](Node_* a, Node_* b) { return a->seq < b->seq; });
  cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
  nodes.reserve(cand.size());
  for (Node_* n : cand) {
    nodes.emplace_back(n, UHM_NOT_SEPARATED);
    n->add_owner(this);
  }
  for (Element_* c : children)
    for (auto& nf : c->nodes)
      if (nf.second == UHM_SEPARATED_SCHUR) nf.first->remove_owner(c);
}

void Element_::restore_connectivity() {
  // element.cxx:362-381 (leaves only)
  if (!is_leaf()) return;
  for (auto& nf : nodes) {
    nf.second = UHM_NOT_SEPARATED;
    nf.first->clean_connectivity();
    nf.first->add_owner(this);
  }
  factor.clear();
  schur.clear();
}

void Element_::numbering(long long& g_offset) {
  // element.cxx:383-392: factor nodes get consecutive global offsets in
  // element-local node order.
  for (auto& nf : nodes)
    if (nf.second == UHM_SEPARATED_FACTOR) {
      nf.first->offset = (int)g_offset;
      g_offset += nf.first->n_dof;
    }
}

static inline long long arrange_key(const Node_* n) {
  // Numbered nodes sort by global offset.  Nodes that are never factored
  // (boundary kind) have no offset in the reference (uninitialised); they are
  // placed after all numbered nodes in creation order.
  return n->offset >= 0 ? (long long)n->offset : (long long)INT_MAX + n->seq;
}

void Element_::arrange_nodes() {
  // element.cxx:288-334: factor and schur lists sorted by global offset with
  // cumulative local offsets; zero-dof nodes are skipped.
  factor.clear();
  schur.clear();
  std::vector<Node_*> f, s;
  for (auto& nf : nodes) {
    if (!nf.first->n_dof) continue;
    if (nf.second == UHM_SEPARATED_FACTOR) f.push_back(nf.first);
    else if (nf.second == UHM_SEPARATED_SCHUR) s.push_back(nf.first);
    else throw std::runtime_error("arrange_nodes: node not separated");
  }
  auto cmp = [](Node_* a, Node_* b) { return arrange_key(a) < arrange_key(b); };
  std::sort(f.begin(), f.end(), cmp);
  std::sort(s.begin(), s.end(), cmp);
  int offs = 0;
  for (Node_* n : f) {
    factor.emplace_back(n, offs);
    offs += n->n_dof;
  }
  fs = offs;
  offs = 0;
  for (Node_* n : s) {
    schur.emplace_back(n, offs);
    offs += n->n_dof;
  }
  ss = offs;
}

std::pair<int, int> Element_::get_n_dof() const {
  std::pair<int, int> r(0, 0);
  for (auto& nf : nodes) {
    if (nf.second == UHM_SEPARATED_FACTOR) r.first += nf.first->n_dof;
    if (nf.second == UHM_SEPARATED_SCHUR) r.second += nf.first->n_dof;
  }
  return r;
}

// LAWN-41 counts, LINAL/src/util.cxx:60-101 (real arithmetic: mul = add = 1).
double flop_gemm(double m, double n, double k) { return 2.0 * m * n * k; }
double flop_syrk(double k, double n) { return 2.0 * (0.5 * k * n * (n + 1.0)); }
double flop_trsm_lower(double m, double n) {
  return 0.5 * n * m * (m + 1.0) + 0.5 * n * m * (m - 1.0);
}
double flop_trsm_upper(double m, double n) {
  return 0.5 * m * n * (n + 1.0) + 0.5 * m * n * (n - 1.0);
}
double flop_lu(double m, double n) {
  if (m > n)
    return (0.5 * m * n * n - (1.0 / 6.0) * n * n * n + 0.5 * m * n - 0.5 * n * n + (2.0 / 3.0) * n) +
           (0.5 * m * n * n - (1.0 / 6.0) * n * n * n - 0.5 * m * n + (1.0 / 6.0) * n);
  return (0.5 * n * m * m - (1.0 / 6.0) * m * m * m + 0.5 * n * m - 0.5 * m * m + (2.0 / 3.0) * m) +
         (0.5 * n * m * m - (1.0 / 6.0) * m * m * m - 0.5 * n * m + (1.0 / 6.0) * m);
}
double flop_chol(double n) {
  return ((1.0 / 6.0) * n * n * n + 0.5 * n * n + (1.0 / 3.0) * n) +
         ((1.0 / 6.0) * n * n * n - (1.0 / 6.0) * n);
}

void Element_::estimate_cost(int method, int, int n_rhs, double& fd, double& fsv, double& nnz,
                             double& buffer) const {
  // element.cxx:128-195
  double f = fs, s = ss, r = n_rhs;
  fsv = flop_trsm_lower(f, r) + flop_gemm(s, r, f) + flop_trsm_upper(r, f) + flop_gemm(f, r, s);
  if (method == UHM_CHOL) {
    fd = flop_chol(f) + flop_trsm_lower(f, s) + flop_syrk(f, s);
    buffer = 8.0 * (f * f + s * f);
    nnz = f * f + s * f;
  } else {
    fd = flop_lu(f, f) + flop_trsm_lower(f, s) + flop_trsm_upper(s, f) + flop_gemm(s, s, f);
    buffer = 8.0 * (f * f + 2.0 * s * f);
    nnz = f * f + 2.0 * s * f;
  }
}

std::vector<Mapper_> build_mapper(const Element_* p, const Element_* c) {
  // helper.hxx:98-158
  auto find_in = [](const std::vector<std::pair<Node_*, int>>& list, const Node_* n) -> int {
    if (n->offset >= 0) {
      auto it = std::lower_bound(list.begin(), list.end(), arrange_key(n),
                                 [](const std::pair<Node_*, int>& a, long long k) { return arrange_key(a.first) < k; });
      if (it != list.end() && it->first == n) return it->second;
      return -1;
    }
    for (auto& e : list)
      if (e.first == n) return e.second;
    return -1;
  };
  std::vector<Mapper_> bij;
  bij.reserve(c->schur.size());
  for (auto& cs : c->schur) {
    Mapper_ q;
    q.offs_c = cs.second;
    int o = find_in(p->factor, cs.first);
    if (o >= 0) {
      q.fs_p = 0;
    } else {
      o = find_in(p->schur, cs.first);
      if (o < 0) throw std::runtime_error("build_mapper: child schur node not found in parent");
      q.fs_p = 1;
    }
    q.offs_p = o;
    q.n_dof = cs.first->n_dof;
    bij.push_back(q);
  }
  std::vector<Mapper_> runs;
  for (size_t i = 0; i < bij.size(); ++i) {
    if (i) {
      int flag = bij[i].fs_p - bij[i - 1].fs_p;
      int diff = bij[i].offs_p - bij[i - 1].offs_p;
      if (!flag && diff == bij[i - 1].n_dof) {
        runs.back().n_dof += bij[i].n_dof;
        continue;
      }
    }
    runs.push_back(bij[i]);
  }
  return runs;
}

// ---------------------------------------------------------------- Mesh_
Mesh_::Mesh_() {}
Mesh_::~Mesh_() {}

Node_* Mesh_::add_node(int id, int phy, int n_dof, int p, int kind) {
  // mesh.cxx:276-295: an existing node is returned when it matches.
  long long k = node_key(id, phy);
  auto it = node_index_.find(k);
  if (it != node_index_.end()) {
    Node_* n = it->second;
    if (n->n_dof != n_dof || n->p != p || n->kind != kind)
      throw std::runtime_error("add_node: node exists with different n_dof/p/kind");
    return n;
  }
  if (locked_) throw std::runtime_error("add_node: mesh is locked");
  nodes_.emplace_back();
  Node_* n = &nodes_.back();
  n->id = id;
  n->phy = phy;
  n->n_dof = n_dof;
  n->p = p;
  n->kind = kind;
  n->seq = (int)nodes_.size() - 1;
  node_index_[k] = n;
  return n;
}

Node_* Mesh_::find_node(int id, int phy) const {
  auto it = node_index_.find(node_key(id, phy));
  return it == node_index_.end() ? nullptr : it->second;
}

void Mesh_::set_node(int id, int phy, int n_dof, int p) {
  if (locked_) throw std::runtime_error("set_node: mesh is locked (unlock first)");
  if (n_dof < 0) throw std::runtime_error("set_node: negative dof count");
  Node_* n = find_node(id, phy);
  if (!n) throw std::runtime_error("set_node: no such node");
  n->n_dof = n_dof;
  n->p = p;
}

Element_* Mesh_::add_element(int gen) {
  if (locked_) throw std::runtime_error("add_element: mesh is locked");
  elements_.emplace_back();
  Element_* e = &elements_.back();
  e->id = id_element_++;
  e->generation = gen;
  return e;
}

Element_* Mesh_::find_element(int id) const {
  if (id < 0 || id >= (int)elements_.size()) return nullptr;
  return const_cast<Element_*>(&elements_[id]);
}

Element_* Mesh_::get_root() const {
  if (elements_.empty()) return nullptr;
  Element_* e = const_cast<Element_*>(&elements_[0]);
  while (e->parent) e = e->parent;
  return e;
}

void Mesh_::update_generation() {
  // Post-order over every tree (op_update_generation, leaf-to-root).
  std::vector<Element_*> order;
  std::vector<std::pair<Element_*, size_t>> st;
  for (auto& e : elements_) {
    if (e.parent) continue;
    st.push_back({&e, 0});
    while (!st.empty()) {
      auto& top = st.back();
      if (top.second < top.first->children.size()) {
        Element_* c = top.first->children[top.second++];
        st.push_back({c, 0});
      } else {
        order.push_back(top.first);
        st.pop_back();
      }
    }
  }
  for (Element_* e : order) {
    if (e->is_leaf()) e->generation = 0;
    else e->update_generation();
  }
}

void Mesh_::lock() {
  // mesh.cxx:438-456
  unlock();
  // The reference relies on build_tree having set generations; a tree whose
  // generations do not decrease towards the root is re-derived here.
  bool ok = true;
  for (auto& e : elements_) {
    if (e.is_leaf() && e.generation != 0) ok = false;
    for (Element_* c : e.children)
      if (c->parent != &e || c->generation <= e.generation) ok = false;
  }
  if (!ok) update_generation();

  // Scheduler_::load (scheduler.cxx:157-178): bucket by generation, element
  // id order inside a generation; leaves in element id order.
  std::map<int, std::vector<Element_*>> buckets;
  leaves_.clear();
  for (auto& e : elements_) {
    buckets[e.generation].push_back(&e);
    if (e.is_leaf()) leaves_.push_back(&e);
  }
  levels_.clear();
  level_gen_.clear();
  for (auto it = buckets.rbegin(); it != buckets.rend(); ++it) {
    levels_.push_back(it->second);
    level_gen_.push_back(it->first);
  }

  for (Element_* e : leaves_) e->restore_connectivity();
  for (auto& lv : levels_)
    for (Element_* e : lv) {
      e->merge_nodes_from_children();
      e->separate_nodes();
    }
  for (auto& n : nodes_) n.offset = -1;
  long long g = 0;
  for (auto& lv : levels_)
    for (Element_* e : lv) e->numbering(g);
  for (auto& lv : levels_)
    for (Element_* e : lv) e->arrange_nodes();
  locked_ = true;
}

void Mesh_::unlock() {
  levels_.clear();
  level_gen_.clear();
  leaves_.clear();
  locked_ = false;
}

void Mesh_::build_tree_coarsening() {
  // build_tree_var_1 (UHM/src/operation/build_tree_var1.cxx:53-181): the
  // METIS-free tree builder by repeated pairwise coarsening of the orphan
  // elements.  The reference orders element pairs and node owners by
  // element address; here element id (creation order) stands in for it, so
  // the tree is deterministic.
  if (locked_) throw std::runtime_error("build_tree: mesh is locked");
  // preprocess (lines 56-61): leaves restore their connectivity, then every
  // element, leaves first, merges its children's schur nodes and separates
  update_generation();
  {
    std::map<int, std::vector<Element_*>> gens;
    for (auto& e : elements_) {
      if (e.is_leaf()) e.restore_connectivity();
      gens[e.generation].push_back(&e);
    }
    for (auto it = gens.rbegin(); it != gens.rend(); ++it)
      for (Element_* e : it->second) {
        e->merge_nodes_from_children();
        e->separate_nodes();
      }
  }
  auto owned_by = [](const Node_* n, const Element_* e) {
    for (const Element_* o : n->owner)
      if (o == e) return true;
    return false;
  };
  // 1. the orphans in element order (lines 74-78)
  std::vector<Element_*> orphan, tmp;
  for (auto& e : elements_)
    if (e.is_orphan()) orphan.push_back(&e);
  // 2. coarsening (lines 81-176)
  while (orphan.size() > 1) {
    // 2.1 branches: pairs of orphans that share a schur node, each with the
    // weight (dofs of the schur nodes the pair shares and nobody else holds,
    // counted once; minus the dofs of every other schur node of either)
    std::set<std::pair<int, int>> branch;
    std::vector<std::pair<std::pair<int, int>, std::pair<Element_*, Element_*>>> weighted;  // insertion order
    for (Element_* o : orphan)
      for (auto& nf : o->nodes) {
        if (nf.second != UHM_SEPARATED_SCHUR) continue;
        std::vector<Element_*> owners = nf.first->owner;
        std::sort(owners.begin(), owners.end(), [](const Element_* a, const Element_* b) { return a->id < b->id; });
        for (Element_* e2 : owners) {
          if (e2 == o) continue;
          Element_* e[2] = {o, e2};
          if (e[0]->id > e[1]->id) std::swap(e[0], e[1]);
          if (!branch.insert({e[0]->id, e[1]->id}).second) continue;
          std::pair<int, int> w(0, 0);
          for (int i = 0; i < 2; ++i)
            for (auto& nn : e[i]->nodes) {
              if (nn.second != UHM_SEPARATED_SCHUR) continue;
              if (nn.first->get_n_owner() == 2 && owned_by(nn.first, e[(i + 1) % 2]))
                w.first += i ? 0 : nn.first->n_dof;
              else
                w.second -= nn.first->n_dof;
            }
          weighted.push_back({w, {e[0], e[1]}});
        }
      }
    // the multimap of the reference: ascending weight, equal weights in
    // insertion order; 2.2 visits it in reverse (largest weight first, equal
    // weights last-inserted first)
    std::stable_sort(weighted.begin(), weighted.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    bool paired = false;
    for (auto it = weighted.rbegin(); it != weighted.rend(); ++it) {
      Element_* e0 = it->second.first;
      Element_* e1 = it->second.second;
      if (!e0->is_orphan() || !e1->is_orphan()) continue;
      const int gen = std::min(e0->generation, e1->generation) - 1;
      Element_* p = add_element(gen);
      for (Element_* c : {e0, e1}) {
        p->add_child(c);
        c->set_parent(p);
      }
      tmp.push_back(p);
      p->merge_nodes_from_children();
      p->separate_nodes();
      paired = true;
    }
    // 2.3 the orphans left over join the new parents
    for (Element_* o : orphan)
      if (o->is_orphan()) tmp.push_back(o);
    std::swap(orphan, tmp);
    tmp.clear();
    // orphans that share no schur node (a disconnected mesh) never pair,
    // and the reference's loop would not end: they are joined by the
    // breadth-first bisection below instead
    if (!paired) break;
  }
  if (orphan.size() > 1) build_tree_bisection();
  update_generation();
}

void Mesh_::build_tree_bisection() {
  // METIS-free fallback for build_tree (the reference's default uses METIS
  // internals, build_tree_var4.cxx:75-252): recursive bisection of the
  // orphan elements in breadth-first order over the shared-node graph.
  std::vector<Element_*> orphans;
  for (auto& e : elements_)
    if (!e.parent) orphans.push_back(&e);
  if (orphans.size() < 2) return;
  std::unordered_map<Node_*, std::vector<int>> owners;
  for (size_t i = 0; i < orphans.size(); ++i)
    for (auto& nf : orphans[i]->nodes) owners[nf.first].push_back((int)i);
  std::vector<int> order, seen(orphans.size(), 0);
  for (size_t s = 0; s < orphans.size(); ++s) {
    if (seen[s]) continue;
    seen[s] = 1;
    size_t head = order.size();
    order.push_back((int)s);
    while (head < order.size()) {
      int i = order[head++];
      for (auto& nf : orphans[i]->nodes)
        for (int j : owners[nf.first])
          if (!seen[j]) {
            seen[j] = 1;
            order.push_back(j);
          }
    }
  }
  std::vector<Element_*> seq;
  for (int i : order) seq.push_back(orphans[i]);
  // recursive halving of the BFS sequence
  struct Rec {
    static Element_* go(Mesh_* m, std::vector<Element_*>& v, size_t lo, size_t hi) {
      if (hi - lo == 1) return v[lo];
      size_t mid = lo + (hi - lo) / 2;
      Element_* a = go(m, v, lo, mid);
      Element_* b = go(m, v, mid, hi);
      Element_* p = m->add_element();
      a->set_parent(p);
      b->set_parent(p);
      p->add_child(a);
      p->add_child(b);
      return p;
    }
  };
  Rec::go(this, seq, 0, seq.size());
  update_generation();
}

void Mesh_::remove_orphan_nodes() {
  // mesh.cxx:246-251.  The kept nodes move to a new container (creation
  // order and seq unchanged) and every element's node pointers follow them.
  if (locked_) throw std::runtime_error("remove_orphan_nodes: mesh is locked");
  std::deque<Node_> kept;
  std::unordered_map<const Node_*, Node_*> to;
  for (Node_& n : nodes_)
    if (n.get_n_owner()) {
      kept.push_back(n);
      to[&n] = &kept.back();
    }
  if (kept.size() == nodes_.size()) return;
  for (Element_& e : elements_) {
    for (auto& nf : e.nodes) nf.first = to.at(nf.first);
    for (auto& nf : e.factor) nf.first = to.at(nf.first);
    for (auto& nf : e.schur) nf.first = to.at(nf.first);
  }
  nodes_.swap(kept);
  node_index_.clear();
  int seq = 0;
  for (Node_& n : nodes_) {
    n.seq = seq++;  // creation indices stay in order, without the gaps
    node_index_[node_key(n.id, n.phy)] = &n;
  }
}

bool Mesh_::import_file(const char* path) {
  // mesh.cxx:511-579: "UHM" / n_nodes / id phy ndof p kind / n_elements /
  // per element: n_nods then one "id phy" line per node; the mesh is emptied
  // first and the nodes no element holds are dropped at the end.
  FILE* fp = fopen(path, "r");
  if (!fp) throw std::runtime_error(std::string("import_file: cannot open ") + path);
  reset();
  char fmt[32] = {0};
  int n_nodes = 0, n_elements = 0;
  auto fail = [&](const char* what) {
    fclose(fp);
    throw std::runtime_error(std::string("import_file: ") + what);
  };
  if (fscanf(fp, "%31s", fmt) != 1 || strcmp(fmt, "UHM")) fail("bad format tag");
  if (fscanf(fp, "%d", &n_nodes) != 1) fail("n_nodes");
  for (int i = 0; i < n_nodes; ++i) {
    int id, phy, ndof, p, kind;
    if (fscanf(fp, "%d %d %d %d %d", &id, &phy, &ndof, &p, &kind) != 5) fail("node line");
    add_node(id, phy, ndof, p, kind);
  }
  if (fscanf(fp, "%d", &n_elements) != 1) fail("n_elements");
  for (int i = 0; i < n_elements; ++i) {
    int nn;
    if (fscanf(fp, "%d", &nn) != 1) fail("element node count");
    Element_* e = add_element();
    for (int j = 0; j < nn; ++j) {
      int id, phy;
      if (fscanf(fp, "%d %d", &id, &phy) != 2) fail("element node line");
      Node_* n = find_node(id, phy);
      if (!n) fail("element references unknown node");
      e->add_node(n);
    }
  }
  fclose(fp);
  remove_orphan_nodes();
  return true;
}

void Mesh_::create_matrix_without_buffer(int datatype, int n_rhs) {
  // mesh/matrix.cxx:64-90: one front per element sized (fs, ss); the
  // fronts of reusable elements keep their factored entries, taken from the
  // previous fronts at the next factorization.
  if (!locked_) throw std::runtime_error("create_matrix_without_buffer: mesh is not locked");
  if (datatype != UHM_REAL) throw std::runtime_error("create_matrix_without_buffer: only UHM_REAL (FLA_DOUBLE) is supported");
  if (n_rhs < 1) throw std::runtime_error("create_matrix_without_buffer: n_rhs < 1");
  std::unique_ptr<FrontSolver> prev = std::move(solver_);
  solver_.reset(new FrontSolver(this, n_rhs));
  if (prev && n_reusable() > 0) solver_->set_reuse_source(std::move(prev));
}

static int reuse_code(const Node_* n) { return n->n_dof + n->p + n->kind; }

void Mesh_::backup(Mesh_& b) const {
  // mesh.cxx:113-150
  if (&b == this) throw std::runtime_error("backup: the backup mesh must be another mesh");
  b.reset();
  for (const Element_& e : elements_) {
    ReuseRecord r;
    for (const auto& nf : e.factor) r.factor.emplace_back(nf.first, reuse_code(nf.first));
    for (const auto& nf : e.schur) r.schur.emplace_back(nf.first, reuse_code(nf.first));
    b.backup_[e.id] = std::move(r);
  }
}

void Mesh_::check_reuse(const Mesh_& b) {
  // mesh.cxx:152-225.  The reference clears the ancestors of a
  // non-reusable element while it walks the elements in id order, so an
  // ancestor visited later gets its own comparison back; here every element
  // is compared first and the ancestors are cleared afterwards, which makes
  // "cannot reuse upper hierarchy" hold whatever the id order.
  auto same = [](const std::vector<std::pair<Node_*, int>>& a, const std::vector<std::pair<const Node_*, int>>& r) {
    if (a.size() != r.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
      if (a[i].first != r[i].first || reuse_code(a[i].first) != r[i].second) return false;
    return true;
  };
  for (Element_& e : elements_) {
    auto it = b.backup_.find(e.id);
    e.reuse = it != b.backup_.end() && same(e.factor, it->second.factor) && same(e.schur, it->second.schur);
  }
  for (Element_& e : elements_)
    if (!e.reuse)
      for (Element_* a = e.parent; a; a = a->parent) a->reuse = false;
}

int Mesh_::n_reusable() const {
  int c = 0;
  for (const Element_& e : elements_) c += e.reuse ? 1 : 0;
  return c;
}

void Mesh_::reset() {
  // mesh.cxx:227-235
  solver_.reset();
  unlock();
  elements_.clear();
  nodes_.clear();
  node_index_.clear();
  backup_.clear();
  id_element_ = 0;
}

void Mesh_::create_matrix_buffer(int) {
  if (!solver_) throw std::runtime_error("create_matrix_buffer: fronts not created");
  solver_->create_buffers();
}

void Mesh_::free_matrix() { solver_.reset(); }

static void phony_child(Mesh_* m, Element_* e, int nod_disp, const int* nods,
                        std::vector<std::pair<Node_*, int>>& list, int& total) {
  // interf.cxx:113-141: the user's node order becomes the schur list of a
  // phony child whose map into the element is the copy-in/out map.
  list.clear();
  total = 0;
  for (size_t i = 0; i < e->nodes.size(); ++i) {
    Node_* n = (nod_disp == UHM_PHYSICS_MULTI) ? m->find_node(nods[2 * i], nods[2 * i + 1])
                                               : m->find_node(nods[i], 0);
    if (!n) throw std::runtime_error("copy_in/out: cannot find node in the mesh");
    if (n->n_dof) {
      list.emplace_back(n, total);
      total += n->n_dof;
    }
  }
}

static std::vector<int> element_map(Element_* e, const std::vector<std::pair<Node_*, int>>& list, int total) {
  Element_ phony;
  phony.schur = list;
  std::vector<Mapper_> runs = build_mapper(e, &phony);
  std::vector<int> map(total);
  for (auto& r : runs)
    for (int d = 0; d < r.n_dof; ++d) map[r.offs_c + d] = (r.fs_p ? e->fs : 0) + r.offs_p + d;
  return map;
}

void Mesh_::copy_in(Element_* e, int datatype, int m, int n, int nod_disp, const int* nods,
                    int side, const double* buffer) {
  // interf.cxx:108-204 (is_in = true)
  if (!solver_) throw std::runtime_error("copy_in: fronts not created");
  if (datatype != UHM_REAL) throw std::runtime_error("copy_in: only UHM_REAL data");
  if (m == 0 || n == 0) throw std::runtime_error("copy_in: dimension is null");
  std::vector<std::pair<Node_*, int>> list;
  int total;
  phony_child(this, e, nod_disp, nods, list, total);
  if (m != total) throw std::runtime_error("copy_in: m does not match the element's dofs");
  std::vector<int> map = element_map(e, list, total);
  if (side == UHM_LHS) solver_->stage_lhs(e->front, m, map.data(), buffer);
  else if (side == UHM_RHS) solver_->stage_rhs(e->front, m, n, map.data(), buffer);
  else throw std::runtime_error("copy_in: side must be UHM_LHS or UHM_RHS");
}

void Mesh_::copy_out(Element_* e, int datatype, int m, int n, int nod_disp, const int* nods,
                     int side, double* buffer) {
  // interf.cxx:108-204 (is_in = false): solution (RHS side) or front entries
  // (LHS side) in the user's node order.
  if (!solver_) throw std::runtime_error("copy_out: fronts not created");
  if (datatype != UHM_REAL) throw std::runtime_error("copy_out: only UHM_REAL data");
  std::vector<std::pair<Node_*, int>> list;
  int total;
  phony_child(this, e, nod_disp, nods, list, total);
  if (m != total) throw std::runtime_error("copy_out: m does not match the element's dofs");
  std::vector<int> map = element_map(e, list, total);
  if (side == UHM_RHS) solver_->fetch_x(e->front, m, n, map.data(), buffer);
  else if (side == UHM_LHS) solver_->fetch_front(e->front, m, map.data(), buffer);
  else throw std::runtime_error("copy_out: side must be UHM_LHS or UHM_RHS");
}

void Mesh_::set_rhs() {
  if (!solver_) throw std::runtime_error("set_rhs: fronts not created");
  solver_->set_rhs();
}

long long Mesh_::get_n_dof() const {
  long long s = 0;
  for (auto& e : elements_) s += e.fs;
  return s;
}

long long Mesh_::get_n_nonzero() const {
  long long s = 0;
  for (auto& e : elements_)
    if (e.is_leaf()) s += (long long)(e.fs + e.ss) * (e.fs + e.ss);
  return s;
}

long long Mesh_::get_n_nonzero_factor() const {
  long long s = 0;
  for (auto& e : elements_) s += (long long)e.fs * e.fs + 2LL * e.fs * e.ss;
  return s;
}

void Mesh_::estimate_cost(int method, int datatype, int n_rhs, double& fd, double& fsv, double& nnz,
                          double& buffer) const {
  fd = fsv = nnz = buffer = 0.0;
  for (auto& e : elements_) {
    double a, b, c, d;
    e.estimate_cost(method, datatype, n_rhs, a, b, c, d);
    fd += a;
    fsv += b;
    nnz += c;
    buffer += d;
  }
}

}  // namespace uhm
