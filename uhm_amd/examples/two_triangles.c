/*
 * A program written against the reference's C/Fortran interface
 * (UHM/src/wrapper/fort.cxx), linked against libuhm_amd.so instead.
 *
 * Two triangles sharing an edge; every triangle has three vertex nodes
 * (1 dof), three edge nodes (3 dofs) and one interior node (3 dofs), 15
 * dofs per element.  The call sequence is the reference's: create the
 * mesh, add elements and nodes, build the tree, lock, create the fronts,
 * copy in the element matrices and right-hand sides, factorize with
 * partial pivoting, solve, check, copy the solution out.
 *
 *   two_triangles            full run on the GPU; prints the residual
 *   two_triangles symbolic   stops after uhm_lock_ (no GPU needed)
 *
 * Exit status 0 on success.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "uhm_amd.h"

#define NDOF_ELT 15
#define NNOD 7

int main(int argc, char** argv) {
  const int symbolic_only = argc > 1 && strcmp(argv[1], "symbolic") == 0;
  /* node ids per element: vertices, edges, interior; the edge 4-5 (node 9)
   * and the vertices 4, 5 are shared */
  static int conn[2][NNOD] = {{3, 4, 5, 7, 9, 8, 1}, {4, 6, 5, 10, 11, 9, 2}};
  static int ndof[NNOD] = {1, 1, 1, 3, 3, 3, 3};

  uhm_fort_p mesh = 0, elt[2], nod;
  uhm_initialize_fla_();
  uhm_mesh_create_(&mesh);
  for (int e = 0; e < 2; ++e) {
    uhm_add_element_(&mesh, &elt[e]);
    for (int i = 0; i < NNOD; ++i) {
      int id = conn[e][i], phy = 0, nd = ndof[i], p = ndof[i];
      uhm_add_node_(&mesh, &id, &phy, &nd, &p, &nod);
      uhm_element_add_node_(&elt[e], &nod);
    }
  }
  uhm_build_tree_(&mesh);
  uhm_lock_(&mesh);
  int n_dof = 0;
  uhm_get_n_dof_(&mesh, &n_dof);
  printf("n_dof %d\n", n_dof);
  if (n_dof != 25) return 1;
  if (symbolic_only) {
    uhm_unlock_(&mesh);
    uhm_mesh_delete_(&mesh);
    return 0;
  }

  int datatype = UHM_AMD_REAL, n_rhs = 1, is_schur = 0;
  uhm_create_matrix_without_buffer_(&mesh, &datatype, &n_rhs);
  uhm_create_element_matrix_buffer_(&mesh, &is_schur);

  /* element matrices: diagonally dominant, unsymmetric, deterministic */
  double A[NDOF_ELT * NDOF_ELT], B[NDOF_ELT];
  unsigned s = 12345u;
  for (int j = 0; j < NDOF_ELT; ++j)
    for (int i = 0; i < NDOF_ELT; ++i) {
      s = s * 1103515245u + 12345u;
      A[i + j * NDOF_ELT] = (double)(s >> 8) / 16777216.0 + (i == j ? NDOF_ELT : 0.0);
    }
  for (int i = 0; i < NDOF_ELT; ++i) B[i] = 1.0 + 0.1 * i;

  int mm = NDOF_ELT, nn = NDOF_ELT, one = 1, disp = UHM_AMD_PHYSICS_SINGLE;
  int lhs = UHM_AMD_LHS, rhs = UHM_AMD_RHS;
  for (int e = 0; e < 2; ++e) {
    uhm_copy_in_(&mesh, &elt[e], &datatype, &mm, &nn, &disp, conn[e], &lhs, A);
    uhm_copy_in_(&mesh, &elt[e], &datatype, &mm, &one, &disp, conn[e], &rhs, B);
  }
  uhm_set_rhs_(&mesh);
  uhm_lu_piv_with_free_(&mesh);
  uhm_solve_lu_piv_(&mesh);
  uhm_check_lu_piv_(&mesh);
  double res = -1.0, eta = -1.0;
  uhm_get_residual_(&mesh, &res);
  uhm_get_backward_error_(&mesh, &eta);
  double X[NDOF_ELT];
  uhm_copy_out_(&mesh, &elt[0], &datatype, &mm, &one, &disp, conn[0], &rhs, X);

  char msg[256];
  if (uhm_get_last_error_(msg, sizeof msg) > 0) {
    fprintf(stderr, "error: %s\n", msg);
    return 2;
  }
  printf("residual %.3e backward_error %.3e x[0] %.15g\n", res, eta, X[0]);
  uhm_unlock_(&mesh);
  uhm_mesh_delete_(&mesh);
  uhm_finalize_fla_();
  return eta <= 1e-14 ? 0 : 3;
}
