/*
 * The two-triangle program of two_triangles.c written against the C++
 * interface (include/uhm_mesh.hpp), in the reference's uhm::Mesh_ style
 * (UHM/include/uhm/mesh/mesh.hxx): nodes and elements through the mesh,
 * build_tree, lock, fronts, copy in, factorize with partial pivoting and
 * with Cholesky on the symmetric part, solve, check, copy out; then the
 * factorization-reuse sequence (backup, change a node, check_reuse).
 *
 *   two_triangles_cpp            full run on the GPU
 *   two_triangles_cpp symbolic   stops after lock (no GPU needed)
 */
#include <cstdio>
#include <cstring>

#include "uhm_mesh.hpp"

int main(int argc, char** argv) {
  const bool symbolic_only = argc > 1 && std::strcmp(argv[1], "symbolic") == 0;
  const int NDOF_ELT = 15, NNOD = 7;
  static int conn[2][7] = {{3, 4, 5, 7, 9, 8, 1}, {4, 6, 5, 10, 11, 9, 2}};
  static const int ndof[7] = {1, 1, 1, 3, 3, 3, 3};
  try {
    uhm::initialize();
    uhm::Mesh m = new uhm::Mesh_;
    uhm::Element elt[2];
    for (int e = 0; e < 2; ++e) {
      elt[e] = m->add_element();
      for (int i = 0; i < NNOD; ++i) elt[e]->add_node(m->add_node(conn[e][i], ndof[i]));
    }
    uhm::build_tree(m);
    m->lock();
    std::printf("n_dof %u\n", m->get_n_dof());
    if (m->get_n_dof() != 25) return 1;
    if (symbolic_only) {
      m->unlock();
      delete m;
      return 0;
    }
    m->create_matrix_without_buffer(uhm::UHM_REAL, 1);
    m->create_element_matrix_buffer();
    double A[15 * 15], S[15 * 15], B[15], X[15];
    unsigned s = 12345u;
    for (int j = 0; j < NDOF_ELT; ++j)
      for (int i = 0; i < NDOF_ELT; ++i) {
        s = s * 1103515245u + 12345u;
        A[i + j * NDOF_ELT] = (double)(s >> 8) / 16777216.0 + (i == j ? NDOF_ELT : 0.0);
      }
    for (int j = 0; j < NDOF_ELT; ++j)
      for (int i = 0; i < NDOF_ELT; ++i) S[i + j * NDOF_ELT] = A[i + j * NDOF_ELT] + A[j + i * NDOF_ELT];
    for (int i = 0; i < NDOF_ELT; ++i) B[i] = 1.0 + 0.1 * i;
    for (int e = 0; e < 2; ++e) {
      m->copy_in(elt[e], uhm::UHM_REAL, NDOF_ELT, NDOF_ELT, conn[e], uhm::UHM_LHS, A);
      m->copy_in(elt[e], uhm::UHM_REAL, NDOF_ELT, 1, conn[e], uhm::UHM_RHS, B);
    }
    m->set_rhs();
    m->lu_piv_with_free();
    m->solve_lu_piv();
    m->check_lu_piv();
    const double res_lu = m->get_residual();
    m->copy_out(elt[0], uhm::UHM_REAL, NDOF_ELT, 1, conn[0], uhm::UHM_RHS, X);
    std::printf("lu residual %.3e x[0] %.15g\n", res_lu, X[0]);

    // Cholesky of the symmetric part, on a fresh set of fronts
    m->create_matrix_without_buffer(uhm::UHM_REAL, 1);
    m->create_element_matrix_buffer();
    for (int e = 0; e < 2; ++e) {
      m->copy_in(elt[e], uhm::UHM_REAL, NDOF_ELT, NDOF_ELT, conn[e], uhm::UHM_LHS, S);
      m->copy_in(elt[e], uhm::UHM_REAL, NDOF_ELT, 1, conn[e], uhm::UHM_RHS, B);
    }
    m->set_rhs();
    m->chol_with_free();
    m->solve_chol();
    m->check_chol();
    const int corrections = m->improve_chol(1e-14, 5);
    const double res_chol = m->get_residual();
    std::printf("chol residual %.3e corrections %d\n", res_chol, corrections);

    // reuse: nothing changed, so every element keeps its front
    uhm::Mesh bk = new uhm::Mesh_;
    m->backup(bk);
    m->check_reuse(bk);
    delete bk;
    m->unlock();
    delete m;
    uhm::finalize();
    const bool ok = res_lu <= 1e-10 && res_chol <= 1e-10;
    std::printf("%s\n", ok ? "ok" : "residual too large");
    return ok ? 0 : 3;
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    return 2;
  }
}
