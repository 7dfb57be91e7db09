/*
 * uhm_amd.h -- drop-in C ABI of the MI355X multifrontal solver.
 *
 * The entry points below carry the names, argument lists and pass-by-pointer
 * conventions of the reference's Fortran/C interface
 * (UHM/include/uhm/wrapper/fort.hxx:38-190, implemented in
 * UHM/src/wrapper/fort.cxx:47-321): handles are `long`, every argument is a
 * pointer, names end with an underscore, functions return void.  A program
 * that binds the reference's interface for the factorize / solve path links
 * against libuhm_amd.so instead and needs no source change.
 *
 * Errors: the reference aborts (assert) or throws; here a failing call
 * leaves its outputs untouched and records a message readable with
 * uhm_get_last_error_ (extension).  A zero pivot or a non-positive Cholesky
 * pivot is reported by uhm_get_info_ (extension) as element id + 1.
 *
 * Scope: the LU-with-partial-pivoting and Cholesky paths.  The reference's
 * lu_nopiv / lu_incpiv / qr / svd entries are not part of this library.
 */
#ifndef UHM_AMD_H_
#define UHM_AMD_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef long uhm_fort_p;          /* fort.hxx:38 */
typedef char uhm_fort_char;
typedef int uhm_fort_int;
typedef unsigned int uhm_fort_uint;
typedef double uhm_fort_double;

/* Constants with the reference's values (UHM/include/uhm/const.hxx:70-81;
 * UHM_REAL is libflame's FLA_DOUBLE code behind LINAL_REAL). */
#define UHM_AMD_REAL 102
#define UHM_AMD_CHOL 1
#define UHM_AMD_LU_PIV 3
#define UHM_AMD_LHS 1
#define UHM_AMD_RHS 2
#define UHM_AMD_PHYSICS_SINGLE 1
#define UHM_AMD_PHYSICS_MULTI 2

/* ---- global state and utilities (fort.hxx:47-74) ------------------------ */
void uhm_initialize_fla_(void);                           /* fort.cxx:47  */
void uhm_finalize_fla_(void);                             /* fort.cxx:51  */
void uhm_set_num_threads_(uhm_fort_int* n_threads);       /* fort.cxx:55  */
void uhm_get_num_threads_(uhm_fort_int* n_threads);       /* fort.cxx:59  */
/* Panel width of the GPU factorization (the reference's hier block size,
 * default 256 at UHM/src/matrix/uhm/matrix.cxx:39); values above 64 are
 * clamped to 64, the widest panel the kernels take. */
void uhm_set_hier_block_size_(uhm_fort_int* size);        /* fort.cxx:63  */
void uhm_get_hier_block_size_(uhm_fort_int* size);        /* fort.cxx:66  */
/* Algorithmic (LAWN-41) flops of the last factorization; the reference's
 * counter is never incremented (UHM/src/matrix/uhm/common.cxx:44-50). */
void uhm_reset_flop_(void);                               /* fort.cxx:69  */
void uhm_get_flop_(uhm_fort_double* flop);                /* fort.cxx:72  */
/* Device bytes held by the library (front arenas, maps, solve and
 * refinement buffers of every mesh of the process; all fronts stay resident
 * in HBM) and their high-water mark, as the reference's process-global
 * buffer_used (UHM/src/matrix/uhm/common.cxx:42-44). */
void uhm_get_buffer_used_(uhm_fort_double* buffer);       /* fort.cxx:75  */
void uhm_get_max_buffer_used_(uhm_fort_double* buffer);   /* fort.cxx:78  */
void uhm_is_multithreading_enable_(uhm_fort_int* flag);   /* fort.cxx:81  */
void uhm_is_hier_matrix_enable_(uhm_fort_int* flag);      /* fort.cxx:84  */
void uhm_timer_(uhm_fort_double* t);                      /* fort.cxx:87  */

/* ---- mesh object (fort.hxx:78-145) --------------------------------------- */
void uhm_mesh_create_(uhm_fort_p* mesh);                                    /* fort.cxx:95  */
void uhm_mesh_delete_(uhm_fort_p* mesh);                                    /* fort.cxx:100 */
void uhm_add_element_(uhm_fort_p* mesh, uhm_fort_p* elt);                   /* fort.cxx:105 */
void uhm_add_node_(uhm_fort_p* mesh, uhm_fort_int* id, uhm_fort_int* phy, uhm_fort_int* n_dof,
                   uhm_fort_int* p, uhm_fort_p* nod);                       /* fort.cxx:111 */
void uhm_element_add_node_(uhm_fort_p* elt, uhm_fort_p* nod);               /* fort.cxx:123 */
void uhm_copy_in_(uhm_fort_p* mesh, uhm_fort_p* elt, uhm_fort_int* datatype, uhm_fort_int* mm,
                  uhm_fort_int* nn, uhm_fort_int* node_disp, uhm_fort_int* nods, uhm_fort_int* side,
                  uhm_fort_double* buffer);                                 /* fort.cxx:130 */
void uhm_copy_out_(uhm_fort_p* mesh, uhm_fort_p* elt, uhm_fort_int* datatype, uhm_fort_int* mm,
                   uhm_fort_int* nn, uhm_fort_int* node_disp, uhm_fort_int* nods, uhm_fort_int* side,
                   uhm_fort_double* buffer);                                /* fort.cxx:145 */
void uhm_mesh_disp_(uhm_fort_p* mesh);                                      /* fort.cxx:160 */
void uhm_get_n_dof_(uhm_fort_p* mesh, uhm_fort_int* n_dof);                 /* fort.cxx:165 */
void uhm_get_n_nonzero_(uhm_fort_p* mesh, uhm_fort_int* n_nonzero);         /* fort.cxx:171 */
void uhm_estimate_cost_(uhm_fort_p* mesh, uhm_fort_int* method, uhm_fort_int* datatype,
                        uhm_fort_int* n_rhs, uhm_fort_double* flop_decompose,
                        uhm_fort_double* flop_solve, uhm_fort_uint* n_nonzero_factor,
                        uhm_fort_double* buffer);                           /* fort.cxx:177 */
void uhm_mesh_import_(uhm_fort_p* mesh, uhm_fort_char* filename);           /* fort.cxx:191 */
/* METIS-free tree: the reference's coarsening builder build_tree_var_1
 * (UHM/src/operation/build_tree_var1.cxx:53-181; element id order stands in
 * for its element-address order), since the reference's default builder
 * needs METIS internals (build_tree_var4.cxx:50-66).  A disconnected mesh's
 * remaining roots are joined by bisecting their breadth-first order. */
void uhm_build_tree_(uhm_fort_p* mesh);                                     /* fort.cxx:197 */
void uhm_create_matrix_without_buffer_(uhm_fort_p* mesh, uhm_fort_int* datatype,
                                       uhm_fort_int* n_rhs);                /* fort.cxx:202 */
void uhm_create_leaf_matrix_buffer_(uhm_fort_p* mesh);                      /* fort.cxx:211 */
void uhm_create_element_matrix_buffer_(uhm_fort_p* mesh, uhm_fort_int* is_schur); /* fort.cxx:216 */
void uhm_set_rhs_(uhm_fort_p* mesh);                                        /* fort.cxx:222 */
/* Element matrices drawn on the device (see uhm_set_random_params_);
 * SPD when beta = 0.  The reference's pair has the SPD flag inverted
 * (UHM/src/mesh/matrix.cxx:155-156, 303-306); here both follow the
 * configured parameters, random_spd_matrix forcing beta = 0. */
void uhm_random_spd_matrix_(uhm_fort_p* mesh);                              /* fort.cxx:227 */
void uhm_random_matrix_(uhm_fort_p* mesh);                                  /* fort.cxx:232 */
/* No effect: the Cholesky kernels never read the upper triangle. */
void uhm_triangularize_(uhm_fort_p* mesh);                                  /* fort.cxx:236 */
/* Factorization reuse after a change of the mesh (no fort.cxx wrapper: the
 * reference offers it through Mesh_::backup / check_reuse,
 * UHM/src/mesh/mesh.cxx:113-225, and the skip in op_decompose,
 * UHM/src/operation/element.cxx:204-206).  uhm_backup_ keeps mesh's
 * connectivity in backup (whose own elements, nodes and fronts are dropped);
 * after the mesh changed (uhm_unlock_, uhm_set_node_, uhm_lock_),
 * uhm_check_reuse_ flags the elements whose factor and schur node lists
 * are unchanged (same nodes, same n_dof + p + kind), none of whose
 * descendants changed; the next uhm_create_matrix_without_buffer_ keeps
 * their factored fronts, which the next factorization takes over and
 * every factorization skips.  n_reusable: flagged elements; n_reused:
 * fronts kept at the last factorization. */
void uhm_set_node_(uhm_fort_p* mesh, uhm_fort_int* id, uhm_fort_int* phy, uhm_fort_int* n_dof, uhm_fort_int* p);
void uhm_backup_(uhm_fort_p* mesh, uhm_fort_p* backup);
void uhm_check_reuse_(uhm_fort_p* mesh, uhm_fort_p* backup);
void uhm_get_n_reusable_(uhm_fort_p* mesh, uhm_fort_int* n);
void uhm_get_n_reused_(uhm_fort_p* mesh, uhm_fort_int* n);
void uhm_lock_(uhm_fort_p* mesh);                                           /* fort.cxx:241 */
void uhm_unlock_(uhm_fort_p* mesh);                                         /* fort.cxx:245 */

/* ---- decomposition / solution / check (fort.hxx:149-186) ---------------- */
void uhm_chol_with_free_(uhm_fort_p* mesh);                                 /* fort.cxx:253 */
void uhm_lu_piv_with_free_(uhm_fort_p* mesh);                               /* fort.cxx:261 */
void uhm_solve_chol_(uhm_fort_p* mesh);                                     /* fort.cxx:277 */
void uhm_solve_lu_piv_(uhm_fort_p* mesh);                                   /* fort.cxx:285 */
/* True residual r = b - A x against the assembled matrix (the reference's
 * hierarchical check multiplies through a no-op trmm,
 * LINAL/src/dense/trmm.cxx:104-130). */
void uhm_check_chol_(uhm_fort_p* mesh);                                     /* fort.cxx:297 */
void uhm_check_lu_piv_(uhm_fort_p* mesh);                                   /* fort.cxx:305 */
/* ||b - A x||_1 after a check (the reference sums ||rt_e||_1 over elements,
 * UHM/src/mesh/matrix.cxx:241-275, which partitions the same rows). */
void uhm_get_residual_(uhm_fort_p* mesh, uhm_fort_double* res);             /* fort.cxx:317 */

/* ---- extensions (no counterpart in fort.hxx) ------------------------------ */
/* Element_::set_parent / add_child (UHM/include/uhm/mesh/element.hxx:97,128),
 * for callers that build the element tree themselves. */
void uhm_element_set_parent_(uhm_fort_p* elt, uhm_fort_p* parent);
void uhm_element_add_child_(uhm_fort_p* elt, uhm_fort_p* child);
/* Synthetic mesh + geometric nested-dissection tree, in place of the
 * ".uhm" files the reference reads (UHM/src/mesh/mesh.cxx:511-579).
 * kind 1: 2D Q_p on dims[0] x dims[1]; 2: 3D Q_p on dims[0..2];
 * 3: 2D variable order p_e in 1..p from seed. */
void uhm_build_synthetic_mesh_(uhm_fort_p* mesh, uhm_fort_int* kind, uhm_fort_int* dims, uhm_fort_int* p,
                               long long* seed);
/* Parameters of uhm_random_*_matrix_: K_e = s_e (M^T M + alpha I + beta S)
 * (alpha < 0 means |alpha| * element size), scale kind 0 none,
 * 1 lognormal(0, p0), 2 10^U(p0, p1); b_e ~ N(0,1). */
void uhm_set_random_params_(uhm_fort_p* mesh, long long* seed, uhm_fort_double* alpha,
                            uhm_fort_double* beta, uhm_fort_int* scale_kind, uhm_fort_double* p0,
                            uhm_fort_double* p1);
/* 0, or element id + 1 of the first front with a zero / non-positive pivot
 * in the last factorization. */
void uhm_get_info_(uhm_fort_p* mesh, uhm_fort_int* info);
/* Normwise backward error ||b - A x||inf / (||A||inf ||x||inf + ||b||inf)
 * of the last check (worst right-hand side). */
void uhm_get_backward_error_(uhm_fort_p* mesh, uhm_fort_double* eta);
/* Milliseconds of the last factorization and solve (device events). */
void uhm_get_timing_(uhm_fort_p* mesh, uhm_fort_double* factor_ms, uhm_fort_double* solve_ms);
/* Per-launch device timing of the next factorizations (on = 1) and the
 * totals of the last profiled one, per kernel kind in the order assemble,
 * lu_panel, lu_rows, chol_diag, chol_trsm, gemm (panel update), schur
 * (final Schur complement), lu_front (whole-front LU of fronts with
 * n <= 64), gemm_diag (Cholesky panel update that also factors the next
 * panel's diagonal blocks), chol_group (a Cholesky panel group in one
 * persistent launch), lp_pack (Schur operands packed to BF16 / FP16),
 * chol_small (assembly and whole Cholesky factorization of the small
 * fronts of a level in one launch), schur_lp (final Schur complement on
 * packed BF16 / FP16 operands), p32_pack (update operands of FP64 fronts
 * copied to FP32), schur_p32 / gemm_p32 (final Schur complement / panel
 * and trailing updates of FP64 fronts with FP32 products on those copies):
 * milliseconds, algorithmic flops, algorithmic bytes and launch counts
 * (arrays of 16). */
void uhm_set_profile_(uhm_fort_p* mesh, uhm_fort_int* on);
void uhm_get_kernel_stats_(uhm_fort_p* mesh, uhm_fort_double* ms, uhm_fort_double* flops, uhm_fort_double* bytes,
                           uhm_fort_int* launches);
/* Precision of the fronts (call after uhm_create_matrix_without_buffer_,
 * before factorizing).  mode 0: every front FP64 with FP64 MFMA updates
 * (default).  mode 1: mixed -- the fronts of the top `top_levels` tree
 * levels stay FP64 (their pivot blocks factorize in FP64), all other fronts
 * are stored in FP32, and every Schur / trailing update is an FP32 MFMA
 * product; the solution is then brought to FP64 accuracy with
 * uhm_improve_*_.  mode 2: as mode 1, but the Schur complement products
 * (ABR -= ABL ATR) take BF16 operands with FP32 accumulation (packed once
 * per front); the panel updates inside each front stay FP32.  mode 3: as
 * mode 2 with FP16 operands, each front's operands scaled by a power of
 * two into FP16's range.  Modes 1-3 need the refinement, modes 2 and 3
 * usually more of it (GMRES-IR, uhm_set_refinement_). */
void uhm_set_precision_(uhm_fort_p* mesh, uhm_fort_int* mode, uhm_fort_int* top_levels);
/* Iterative refinement after uhm_solve_*_: x <- x + M^-1 (b - A x), the
 * residual in FP64 against the assembled matrix, M the last factorization,
 * until the normwise backward error of every right-hand side is <= tol or
 * max_iter corrections were applied; n_iter = corrections applied.  Fills
 * the slot of Mesh_::improve_chol / improve_lu_piv
 * (UHM/src/mesh/chol.cxx:136-149, UHM/src/mesh/lu_piv.cxx:232-245), whose
 * reference bodies are stubs. */
void uhm_improve_chol_(uhm_fort_p* mesh, uhm_fort_double* tol, uhm_fort_int* max_iter, uhm_fort_int* n_iter);
void uhm_improve_lu_piv_(uhm_fort_p* mesh, uhm_fort_double* tol, uhm_fort_int* max_iter, uhm_fort_int* n_iter);
/* Backward error before each correction of the last refinement (n in:
 * capacity of eta, out: number of entries) and its wall time in ms. */
void uhm_get_refine_history_(uhm_fort_p* mesh, uhm_fort_double* eta, uhm_fort_int* n, uhm_fort_double* ms);
/* Correction kind of uhm_improve_*_: kind 0 classical (d = M^-1 r,
 * default), 1 GMRES-IR (d from restarted GMRES(restart) on the
 * left-preconditioned system M^-1 A d = M^-1 r, stopped when the
 * preconditioned residual falls to inner_tol times its start; restart
 * 1..31), 2 classical while every correction cuts the backward error at
 * least tenfold, GMRES-IR from the first one that does not.  Extension
 * (no reference counterpart: the reference's improve_* are stubs). */
void uhm_set_refinement_(uhm_fort_p* mesh, uhm_fort_int* kind, uhm_fort_int* restart, uhm_fort_double* inner_tol);
/* Applications of M^-1 (level solves) in the last refinement. */
void uhm_get_refine_solves_(uhm_fort_p* mesh, uhm_fort_int* n);
/* Schedule knobs: key 1 = Cholesky panels per trailing update (default 0:
 * 6), key 2 = mean tile extent from
 * which a GEMM launch uses 128 x 128 tiles (default 8192: 64 x 64 tiles on
 * every workload measured, whose 128-tile launches ran slower), key 3 = number
 * of right-hand sides from which the solves run as block solves, FP64 MFMA
 * products over 64-column tiles (default 8), key 4 = largest LU front
 * (n <= 64) factored by the one-launch whole-front kernel (default 64; 0:
 * every front through the panel path), key 5 = Cholesky panel lookahead:
 * the next panel's diagonal block and TRSM run on a second stream beside
 * the current panel's trailing update (default 0: off; the stream
 * dependencies cost more than the overlap wins on the measured meshes, the
 * autotuner tries it), key 6 = Cholesky diagonal blocks inside a panel
 * group factored by the workgroup of the previous panel update that
 * updates them, without a launch of their own (default 1; 0: one diagonal
 * launch per panel; the factors are the same bits either way), key 7 =
 * Cholesky panel groups (key 1, at most 16 panels) each in one persistent
 * launch whose workgroups resolve the panel-to-panel dependencies among
 * themselves, left-looking inside the group: 0 never (the per-panel
 * launches that keys 5 and 6 shape), 1 on every tree level, 2 (default) on
 * the levels with at most 16 fronts, where they measured faster; key 8 =
 * Cholesky level solves (1, default: one launch per tree level and
 * direction for up to 8 right-hand sides; 0: per-block launches); key 9 =
 * Cholesky tree levels of at least this many fronts, every one with
 * fs <= 64, assembled and factored in one launch, a workgroup per front,
 * with the same factor bits as the four level launches (default 0: never;
 * on the 16^3 Q4 mesh the fused launch ran its levels 1.2-1.4x slower than
 * the four launches, whose separate kernels keep 4-8 workgroups per CU
 * where the fused one holds 2); key 10 = precision modes 2 and 3: the
 * smallest mean Schur block (rows) of a tree level whose Schur products
 * take packed BF16 / FP16 operands, the levels below it FP32 products
 * (default 768: on small blocks the packing saves next to nothing and the
 * rounding costs refinement steps; 0: every level); keys 100 + c,
 * 110 + c, 120 + c = the panel group, tile threshold and group-launch
 * policy (as keys 1, 2, 7) of the fronts of size class c (c = 0: fs <= 64,
 * 1: fs <= 512, 2: larger), whose fronts each tree level schedules as
 * separate launches, class after class (0 / 0 / -1, the defaults: the
 * global key). */
void uhm_set_tuning_(uhm_fort_p* mesh, uhm_fort_int* key, uhm_fort_int* value);
void uhm_get_tuning_(uhm_fort_p* mesh, uhm_fort_int* key, uhm_fort_int* value);
/* Autotuner: times factorizations of the current values over the knobs
 * (coordinate search: keys 1, 2, 5, then keys 100 + c and 110 + c of every
 * size class the pattern has) and keeps the fastest setting; best_ms = its
 * time. */
void uhm_autotune_(uhm_fort_p* mesh, uhm_fort_int* method, uhm_fort_double* best_ms);
/* Graph mode (on = 1): the factorization's and each solve's launch
 * sequence is captured once as a HIP graph and replayed afterwards, for
 * re-factorization of new element values on a fixed sparsity pattern. */
void uhm_set_graph_(uhm_fort_p* mesh, uhm_fort_int* on);
/* Last error message (empty when none); returns its length. */
int uhm_get_last_error_(char* buf, int len);

/* Introspection of the symbolic analysis and of device results (tests). */
void uhm_get_n_fronts_(uhm_fort_p* mesh, uhm_fort_int* n);
/* out[0..8] = element id, fs, ss, level, parent front, generation,
 * factor nodes, schur nodes, is_leaf. */
void uhm_get_front_info_(uhm_fort_p* mesh, uhm_fort_int* front, uhm_fort_int* out);
/* which 0: factor list, 1: schur list; node ids and local offsets
 * (Element_::factor / schur, UHM/include/uhm/mesh/element.hxx:61). */
void uhm_get_front_nodes_(uhm_fort_p* mesh, uhm_fort_int* front, uhm_fort_int* which, uhm_fort_int* ids,
                          uhm_fort_int* offs, uhm_fort_int* n);
/* Extend-add runs of (parent, child front) as {offs_c, offs_p, fs_p, n_dof}
 * (Helper_::set_mapper, UHM/include/uhm/matrix/uhm/helper.hxx:98-158). */
void uhm_get_mapper_(uhm_fort_p* mesh, uhm_fort_int* child_front, uhm_fort_int* runs, uhm_fort_int* n);
void uhm_get_node_offset_(uhm_fort_p* mesh, uhm_fort_int* id, uhm_fort_int* phy, uhm_fort_int* offset);
void uhm_fetch_front_(uhm_fort_p* mesh, uhm_fort_int* front, uhm_fort_double* out);
void uhm_fetch_pivots_(uhm_fort_p* mesh, uhm_fort_int* front, uhm_fort_int* out);
void uhm_fetch_stage_(uhm_fort_p* mesh, uhm_fort_int* front, uhm_fort_double* out);
/* right-hand side block b of a front (n x n_rhs, front order) */
void uhm_fetch_rhs_front_(uhm_fort_p* mesh, uhm_fort_int* front, uhm_fort_double* out);
void uhm_fetch_x_global_(uhm_fort_p* mesh, uhm_fort_double* out);
void uhm_fetch_b_global_(uhm_fort_p* mesh, uhm_fort_double* out);
/* Connectivity dump (the role of Mesh_::export_connectivity,
 * UHM/include/uhm/mesh/mesh.hxx:93): nodes in creation order (id, n_dof,
 * kind), element node lists as creation indices (eptr/enode), parent
 * element index and children lists (cptr/cidx).  Call with null arrays to
 * get the sizes (n_nodes, n_elem, len(enode), len(cidx)) in sizes[0..3]. */
void uhm_export_connectivity_(uhm_fort_p* mesh, uhm_fort_int* sizes, uhm_fort_int* node_id, uhm_fort_int* ndof,
                              uhm_fort_int* kind, uhm_fort_int* eptr, uhm_fort_int* enode, uhm_fort_int* parent,
                              uhm_fort_int* cptr, uhm_fort_int* cidx);
void uhm_get_csr_size_(uhm_fort_p* mesh, long long* nnz);
void uhm_fetch_csr_(uhm_fort_p* mesh, long long* rowptr, uhm_fort_int* col, uhm_fort_double* val);

#ifdef __cplusplus
}
#endif
#endif
