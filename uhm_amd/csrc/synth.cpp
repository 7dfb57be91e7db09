// Synthetic finite-element meshes and their geometric nested-dissection
// element trees, built only through the public mesh API (add_node,
// add_element, Element_::add_node / set_parent / add_child), the way a user
// of the reference drives Mesh_ (UHM/example/cxx/ex1.cxx:16-101).  They
// replace the ".uhm" input files, which the reference does not ship.
//
// Uniform Q_p on an N0 x N1 (x N2) box: one node (one dof) per Lagrange
// point, node id = linear grid index (x fastest), elements in lexicographic
// order (x fastest), element node order = tensor order (x fastest).
//
// Variable-p 2D: per-element order p_e in {1..pmax} drawn from Philox
// (seed, element); entity nodes with dofs vertex 1, edge min(p) - 1 (zero is
// allowed and kept), interior (p_e - 1)^2.
//
// Tree: recursive bisection of the element box along its longest axis (the
// first on ties) at lo + len/2; parents are created in post order.
#include <algorithm>
#include <array>
#include <stdexcept>
#include <vector>

#include "mesh.hpp"

namespace uhm {

namespace {

// Philox-4x32-10, same constants as the device generator.
inline void philox(unsigned c[4], unsigned k0, unsigned k1) {
  for (int r = 0; r < 10; ++r) {
    const unsigned long long p0 = (unsigned long long)0xD2511F53u * c[0];
    const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c[2];
    const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
    const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
    const unsigned n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

Element_* nd_tree(Mesh_* m, const std::vector<Element_*>& cells, std::array<int, 3> N, int d,
                  std::array<int, 3> lo, std::array<int, 3> hi) {
  int axis = 0, best = -1;
  for (int a = 0; a < d; ++a)
    if (hi[a] - lo[a] > best) {
      best = hi[a] - lo[a];
      axis = a;
    }
  if (best == 1) {
    const int idx = lo[0] + N[0] * (lo[1] + N[1] * (d == 3 ? lo[2] : 0));
    return cells[idx];
  }
  const int mid = lo[axis] + (hi[axis] - lo[axis]) / 2;
  std::array<int, 3> hiL = hi, loR = lo;
  hiL[axis] = mid;
  loR[axis] = mid;
  Element_* a = nd_tree(m, cells, N, d, lo, hiL);
  Element_* b = nd_tree(m, cells, N, d, loR, hi);
  Element_* p = m->add_element();
  a->set_parent(p);
  b->set_parent(p);
  p->add_child(a);
  p->add_child(b);
  return p;
}

}  // namespace

int synth_p_of(unsigned long long seed, int elem, int pmax) {
  unsigned c[4] = {(unsigned)elem, 0x7u, 0u, 0u};
  philox(c, (unsigned)seed, (unsigned)(seed >> 32));
  return 1 + (int)(c[0] % (unsigned)pmax);
}

// kind 1: uniform Q_p 2D; 2: uniform Q_p 3D; 3: variable-p 2D (p in 1..pmax).
void build_synthetic_mesh(Mesh_* m, int kind, const int* dims, int p, unsigned long long seed) {
  if (m->get_n_elements() || m->get_n_nodes()) throw std::runtime_error("synthetic mesh: mesh is not empty");
  const int d = (kind == 2) ? 3 : 2;
  std::array<int, 3> N = {dims[0], dims[1], d == 3 ? dims[2] : 1};
  for (int a = 0; a < d; ++a)
    if (N[a] < 1) throw std::runtime_error("synthetic mesh: bad dimensions");
  if (p < 1) throw std::runtime_error("synthetic mesh: p < 1");
  std::vector<Element_*> cells;
  if (kind == 1 || kind == 2) {
    const int X = N[0] * p + 1, Y = N[1] * p + 1, Z = (d == 3) ? N[2] * p + 1 : 1;
    const long long nn = (long long)X * Y * Z;
    if (nn > 0x7fffffffLL) throw std::runtime_error("synthetic mesh: too many nodes");
    std::vector<Node_*> nodes(nn);
    for (long long i = 0; i < nn; ++i) nodes[i] = m->add_node((int)i, 0, 1, p, UHM_NODE_KIND_DEFAULT);
    const int Pz = (d == 3) ? p : 0;
    for (int ez = 0; ez < N[2]; ++ez)
      for (int ey = 0; ey < N[1]; ++ey)
        for (int ex = 0; ex < N[0]; ++ex) {
          Element_* e = m->add_element();
          for (int c = 0; c <= Pz; ++c)
            for (int b = 0; b <= p; ++b)
              for (int a = 0; a <= p; ++a) {
                const long long id = (ex * p + a) + (long long)X * ((ey * p + b) + (long long)Y * (ez * p + c));
                e->add_node(nodes[id]);
              }
          cells.push_back(e);
        }
  } else if (kind == 3) {
    const int N0 = N[0], N1 = N[1];
    std::vector<int> pe(N0 * N1);
    for (int e = 0; e < N0 * N1; ++e) pe[e] = synth_p_of(seed, e, p);
    auto P = [&](int i, int j) { return pe[i + N0 * j]; };
    int id = 0;
    std::vector<Node_*> vert((N0 + 1) * (N1 + 1)), xe(N0 * (N1 + 1)), ye((N0 + 1) * N1), in(N0 * N1);
    for (int j = 0; j <= N1; ++j)
      for (int i = 0; i <= N0; ++i) vert[i + (N0 + 1) * j] = m->add_node(id++, 0, 1, 1, UHM_NODE_KIND_DEFAULT);
    for (int j = 0; j <= N1; ++j)
      for (int i = 0; i < N0; ++i) {
        int pm = 1 << 30;
        if (j > 0) pm = std::min(pm, P(i, j - 1));
        if (j < N1) pm = std::min(pm, P(i, j));
        xe[i + N0 * j] = m->add_node(id++, 0, pm - 1, pm, UHM_NODE_KIND_DEFAULT);
      }
    for (int j = 0; j < N1; ++j)
      for (int i = 0; i <= N0; ++i) {
        int pm = 1 << 30;
        if (i > 0) pm = std::min(pm, P(i - 1, j));
        if (i < N0) pm = std::min(pm, P(i, j));
        ye[i + (N0 + 1) * j] = m->add_node(id++, 0, pm - 1, pm, UHM_NODE_KIND_DEFAULT);
      }
    for (int j = 0; j < N1; ++j)
      for (int i = 0; i < N0; ++i) {
        const int q = P(i, j);
        in[i + N0 * j] = m->add_node(id++, 0, (q - 1) * (q - 1), q, UHM_NODE_KIND_DEFAULT);
      }
    for (int j = 0; j < N1; ++j)
      for (int i = 0; i < N0; ++i) {
        Element_* e = m->add_element();
        e->add_node(vert[i + (N0 + 1) * j]);
        e->add_node(vert[i + 1 + (N0 + 1) * j]);
        e->add_node(vert[i + (N0 + 1) * (j + 1)]);
        e->add_node(vert[i + 1 + (N0 + 1) * (j + 1)]);
        e->add_node(xe[i + N0 * j]);
        e->add_node(xe[i + N0 * (j + 1)]);
        e->add_node(ye[i + (N0 + 1) * j]);
        e->add_node(ye[i + 1 + (N0 + 1) * j]);
        e->add_node(in[i + N0 * j]);
        cells.push_back(e);
      }
  } else {
    throw std::runtime_error("synthetic mesh: unknown kind");
  }
  if (cells.size() > 1) nd_tree(m, cells, N, d, {0, 0, 0}, {N[0], N[1], d == 3 ? N[2] : 1});
  m->update_generation();
}

}  // namespace uhm
