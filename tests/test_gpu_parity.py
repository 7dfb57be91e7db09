"""GPU parity: the HIP path through the C ABI against the CPU oracle and the
golden vectors.  Bars: integer / index work (pivots, assembly) bit-exact;
FP64 factors within 1e-10 of the oracle's (max-norm relative to the front);
solutions within 1e-12 relative of the oracle's and the golden dense solve;
normwise backward error <= 1e-13 on the assembled matrix."""
import ctypes as C
import sys

import numpy as np
import pytest

import uhm_amd as U
from meshes import load_q2, qp_spec, spec_from_connectivity

pytestmark = pytest.mark.gpu

FACTOR_TOL = 1e-10
X_TOL = 1e-12
ETA_TOL = 1e-13


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _copy_in_golden(m, E, spec, z, N):
    el, K, b = z["N%d_elems" % N], z["N%d_K" % N], z["N%d_b" % N]
    for e in range(len(el)):
        m.copy_in(E[e], el[e], U.UHM_LHS, K[e])
        m.copy_in(E[e], el[e], U.UHM_RHS, b[e].reshape(-1, 1))


def _oracle_from_product(m, method, b=64):
    """Oracle on the same mesh and the same input bits (the device stages)."""
    spec = spec_from_connectivity(m.export_connectivity())
    om = spec.oracle()
    om.init_numeric(m.n_rhs)
    for f in range(m.n_fronts()):
        info = m.front_info(f)
        if info["leaf"]:
            om.set_front(info["id"], m.fetch_stage(f))
    return spec, om


def _compare_fronts(m, om, method):
    worst = 0.0
    for f in range(m.n_fronts()):
        info = m.front_info(f)
        e, fs = info["id"], info["fs"]
        G = m.fetch_front(f)
        R = om.front(e)
        if method == U.UHM_CHOL:
            lo = np.tril(np.ones(G.shape, bool))
            G, R = G[lo], R[lo]
        else:
            pg, pr = m.fetch_pivots(f), om.pivots(e)
            assert np.array_equal(pg, pr), "pivot mismatch in front %d (element %d)" % (f, e)
        if G.size:
            err = np.abs(G - R).max() / max(1e-300, np.abs(R).max())
            worst = max(worst, err)
            assert err <= FACTOR_TOL, (f, e, err)
    return worst


def kappa_1(m):
    """1-norm condition number of the assembled matrix (scipy splu +
    onenormest)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    rp, col, val = m.fetch_csr()
    n = len(rp) - 1
    A = sp.csr_matrix((val, col, rp), shape=(n, n)).tocsc()
    lu = spl.splu(A)
    op = spl.LinearOperator((n, n), matvec=lu.solve, rmatvec=lambda v: lu.solve(v, trans="T"), dtype=float)
    return spl.onenormest(A) * spl.onenormest(op)


def x_tolerance(m):
    """Forward-error bar for two backward-stable FP64 solutions of the same
    system: max(1e-12, 10 kappa_1(A) eps)."""
    kappa = kappa_1(m)
    tol = max(X_TOL, 10.0 * kappa * np.finfo(float).eps)
    say("  kappa_1 ~ %.3e -> x tolerance %.2e" % (kappa, tol))
    return tol


def _solution_by_node(m, x):
    spec = spec_from_connectivity(m.export_connectivity())
    out = np.zeros(len(spec.nodes))
    for nid, _ in spec.nodes:
        out[nid] = x[m.node_offset(nid)]
    return out


@pytest.mark.parametrize("N,method", [(4, U.UHM_LU_PIV), (16, U.UHM_CHOL)])
def test_golden_q2_meshes(lib, oracle_lib, N, method):
    """Copy-in path on the golden Q2 meshes (16x16 is the 1089-dof CPU
    configuration): solution against numpy.linalg.solve of the assembled
    system, every front against the oracle."""
    say("case golden q2 N=%d method=%d" % (N, method))
    z = load_q2()
    spec = qp_spec((N, N), 2)
    m, E = spec.build_product(U)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    _copy_in_golden(m, E, spec, z, N)
    m.set_rhs()
    assert (m.chol_with_free() if method == U.UHM_CHOL else m.lu_piv_with_free()) == 0
    (m.solve_chol if method == U.UHM_CHOL else m.solve_lu_piv)()
    (m.check_chol if method == U.UHM_CHOL else m.check_lu_piv)()
    x = _solution_by_node(m, m.fetch_x()[:, 0])
    ref = z["N%d_x" % N]
    assert np.abs(x - ref).max() / np.abs(ref).max() <= X_TOL
    assert m.backward_error() <= ETA_TOL
    # copy_out returns the element's solution in the user's node order
    xe = m.copy_out(E[0], spec.elems[0], U.UHM_RHS, len(spec.elems[0]), 1)[:, 0]
    np.testing.assert_allclose(xe, ref[spec.elems[0]], rtol=0, atol=X_TOL * np.abs(ref).max())
    # oracle on the same input
    om = spec.oracle()
    om.init_numeric(1)
    el, K, b = z["N%d_elems" % N], z["N%d_K" % N], z["N%d_b" % N]
    for e in range(len(el)):
        om.copy_in(e, [spec.seq[n] for n in el[e]], 1, K[e])
    assert om.factor(method, 64) == 0
    _compare_fronts(m, om, method)


CASES = [
    # name, kind, dims, p, method, alpha, beta, scale_kind, p0, p1
    ("q3-3d-lu", 2, (4, 4, 4), 3, U.UHM_LU_PIV, 64.0, 0.1, 1, 1.0, 0.0),
    ("q4-3d-chol", 2, (4, 4, 3), 4, U.UHM_CHOL, 125.0, 0.0, 2, -2.0, 2.0),
    ("q3-3d-chol-deep", 2, (6, 6, 6), 3, U.UHM_CHOL, 64.0, 0.0, 1, 0.5, 0.0),
    ("q3-3d-lu-deep", 2, (6, 6, 5), 3, U.UHM_LU_PIV, 64.0, 0.1, 1, 1.0, 0.0),
    ("varp-2d-lu", 3, (24, 24), 8, U.UHM_LU_PIV, -1.0, 0.1, 0, 0.0, 0.0),
    ("q2-2d-chol", 1, (20, 12), 2, U.UHM_CHOL, 9.0, 0.0, 0, 0.0, 0.0),
    # fronts above 768 rows: multi-panel pivoting (128-wide GEMM tiles: the
    # tile-threshold knob in test_schedule_knobs_preserve_results)
    ("q3-3d-chol-8cube", 2, (8, 8, 8), 3, U.UHM_CHOL, 64.0, 0.0, 1, 0.5, 0.0),
    ("q3-3d-lu-8cube", 2, (8, 8, 8), 3, U.UHM_LU_PIV, 64.0, 0.1, 1, 1.0, 0.0),
    # root fronts above 1024 rows: super-block triangular solves
    ("q4-3d-chol-8cube", 2, (8, 8, 8), 4, U.UHM_CHOL, 125.0, 0.0, 2, -2.0, 2.0),
    ("q3-3d-lu-11cube", 2, (11, 11, 11), 3, U.UHM_LU_PIV, 64.0, 0.1, 1, 1.0, 0.0),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_device_generated_vs_oracle(lib, oracle_lib, case):
    name, kind, dims, p, method, alpha, beta, sk, p0, p1 = case
    say("case", name)
    m = U.Mesh()
    m.build_synthetic(kind, dims, p, 12345)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=7, alpha=alpha, beta=beta, scale_kind=sk, p0=p0, p1=p1)
    m.random_matrix()
    m.set_rhs()
    chol = method == U.UHM_CHOL
    assert (m.chol_with_free() if chol else m.lu_piv_with_free()) == 0
    (m.solve_chol if chol else m.solve_lu_piv)()
    (m.check_chol if chol else m.check_lu_piv)()
    eta = m.backward_error()
    say("  eta", eta)
    assert eta <= ETA_TOL
    spec, om = _oracle_from_product(m, method)
    for f in range(m.n_fronts()):
        info = m.front_info(f)
        if info["leaf"]:
            om.set_rhs_front(info["id"], m.fetch_rhs_front(f))
    assert om.factor(method, 64) == 0
    worst = _compare_fronts(m, om, method)
    say("  worst front rel err", worst)
    om.solve(method)
    xo = om.x_global(m.get_n_dof())[:, 0]
    xg = m.fetch_x()[:, 0]
    err = np.abs(xg - xo).max() / np.abs(xo).max()
    say("  x rel err vs oracle", err)
    assert err <= x_tolerance(m)


def test_solution_matches_oracle_device_rhs(lib, oracle_lib):
    """Same input bits on both sides including the right-hand side: the GPU
    solution equals the oracle's to 1e-12 relative."""
    m = U.Mesh()
    m.build_synthetic(2, (5, 4, 4), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=3, alpha=64.0, beta=0.1, scale_kind=1, p0=1.0)
    m.random_matrix()
    m.set_rhs()
    assert m.lu_piv_with_free() == 0
    m.solve_lu_piv()
    xg = m.fetch_x()[:, 0]
    spec, om = _oracle_from_product(m, U.UHM_LU_PIV)
    for f in range(m.n_fronts()):
        info = m.front_info(f)
        if info["leaf"]:
            om.set_rhs_front(info["id"], m.fetch_rhs_front(f))
    assert om.factor(U.UHM_LU_PIV, 64) == 0
    om.solve(U.UHM_LU_PIV)
    xo = om.x_global(len(xg))[:, 0]
    err = np.abs(xg - xo).max() / np.abs(xo).max()
    say("  x rel err vs oracle", err)
    assert err <= X_TOL
    # a second solve on the same factors reproduces the bits
    m.solve_lu_piv()
    assert np.array_equal(m.fetch_x()[:, 0], xg)


@pytest.mark.parametrize("method", [U.UHM_LU_PIV, U.UHM_CHOL])
def test_multiple_rhs_vs_oracle(lib, oracle_lib, method):
    """Three right-hand sides solved together match the oracle column by
    column, and each column's backward error is at FP64 level."""
    m = U.Mesh()
    m.build_synthetic(2, (5, 4, 4), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 3)
    m.create_matrix_buffer()
    m.set_random_params(seed=5, alpha=64.0, beta=0.1 if method == U.UHM_LU_PIV else 0.0, scale_kind=1, p0=0.5)
    m.random_matrix()
    m.set_rhs()
    chol = method == U.UHM_CHOL
    assert (m.chol_with_free() if chol else m.lu_piv_with_free()) == 0
    (m.solve_chol if chol else m.solve_lu_piv)()
    (m.check_chol if chol else m.check_lu_piv)()
    assert m.backward_error() <= ETA_TOL
    xg = m.fetch_x()
    assert xg.shape[1] == 3
    spec, om = _oracle_from_product(m, method)
    for f in range(m.n_fronts()):
        info = m.front_info(f)
        if info["leaf"]:
            om.set_rhs_front(info["id"], m.fetch_rhs_front(f))
    assert om.factor(method, 64) == 0
    om.solve(method)
    xo = om.x_global(m.get_n_dof())
    for c in range(3):
        err = np.abs(xg[:, c] - xo[:, c]).max() / np.abs(xo[:, c]).max()
        say("  rhs", c, "x rel err vs oracle", err)
        assert err <= X_TOL
    assert not np.array_equal(xg[:, 0], xg[:, 1])


def test_refactorization_reproducible(lib):
    """A second factorization of the same values gives the same bits."""
    m = U.Mesh()
    m.build_synthetic(2, (4, 4, 4), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=11, alpha=64.0, beta=0.1, scale_kind=1, p0=1.0)
    m.random_matrix()
    m.set_rhs()
    assert m.lu_piv_with_free() == 0
    F1 = [m.fetch_front(f) for f in range(m.n_fronts())]
    assert m.lu_piv_with_free() == 0
    for f in range(m.n_fronts()):
        assert np.array_equal(F1[f], m.fetch_front(f))


def test_zero_pivot_reports_element(lib):
    """A singular leaf ATL is reported (element id + 1), not aborted."""
    spec = qp_spec((2, 2), 1)
    m, E = spec.build_product(U)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    rng = np.random.default_rng(0)
    for e in range(4):
        K = rng.standard_normal((4, 4)) + 4 * np.eye(4)
        if e == 0:
            K[:] = 0.0
        m.copy_in(E[e], spec.elems[e], U.UHM_LHS, K)
    info = m.lu_piv_with_free()
    assert info >= 1


def test_schedule_knobs_preserve_results(lib):
    """Every Cholesky panel grouping and tile threshold the autotuner may
    pick factors the same matrix to rounding (different summation order
    only), and the autotuner leaves a valid setting behind."""
    m = U.Mesh()
    m.build_synthetic(2, (8, 8, 8), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=4, alpha=64.0, scale_kind=1, p0=0.5)
    m.random_matrix()
    m.set_rhs()
    assert m.chol_with_free() == 0
    ref = [np.tril(m.fetch_front(f)) for f in range(m.n_fronts())]
    # global knobs, then the per-size-class ones (class 1: 64 < fs <= 512,
    # class 2: larger; keys 100 + c panel group, 110 + c tile threshold,
    # 120 + c group launches)
    knobs = [(1, 1), (1, 2), (1, 8), (1, 12), (2, 64), (2, 1024),
             (101, 2), (102, 1), (111, 64), (112, 64), (121, 0), (122, 1)]
    defaults = {1: 0, 2: 8192, 101: 0, 102: 0, 111: 0, 112: 0, 121: -1, 122: -1}
    for key, val in knobs:
        m.set_tuning(key, val)
        assert m.get_tuning(key) == val
        assert m.chol_with_free() == 0
        for f in range(m.n_fronts()):
            G = np.tril(m.fetch_front(f))
            err = np.abs(G - ref[f]).max() / max(1e-300, np.abs(ref[f]).max())
            assert err <= 1e-12, (key, val, f, err)
        m.set_tuning(key, defaults[key])
    classes = sorted({0 if fs <= 64 else (1 if fs <= 512 else 2)
                      for fs in (m.front_info(f)["fs"] for f in range(m.n_fronts())) if fs > 0})
    assert len(classes) >= 2
    best, knobs = m.autotune(U.UHM_CHOL)
    say("  autotune", best, knobs)
    assert best > 0 and knobs["chol_group"] in (1, 2, 4, 6, 8, 12)
    for c in classes:
        assert knobs["class%d" % c]["tile_threshold"] in (0, 512, 2048, 8192)
        if c > 0:
            assert knobs["class%d" % c]["chol_group"] in (0, 1, 2, 4, 6, 8, 12)
    assert m.chol_with_free() == 0
    m.solve_chol()
    m.check_chol()
    assert m.backward_error() <= ETA_TOL


@pytest.mark.parametrize("dims,p,prec", [((4, 4, 4), 3, U.PREC_FP64), ((5, 4, 3), 2, U.PREC_FP64),
                                         ((4, 4, 4), 3, U.PREC_MIXED)])
def test_lu_whole_front_kernel_matches_panel_path(lib, dims, p, prec):
    """Fronts with n <= 64 factored by the whole-front kernel give the
    panel path's pivots exactly and its factors to rounding (the Schur
    update's summation order differs: rank-1 FMAs against MFMA tiles)."""
    m = U.Mesh()
    m.build_synthetic(2, dims, p, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=13, alpha=64.0, beta=0.1, scale_kind=1, p0=1.0)
    m.random_matrix()
    m.set_rhs()
    m.set_precision(prec, 1)
    m.set_tuning(4, 0)
    assert m.lu_piv_with_free() == 0
    F0 = [m.fetch_front(f) for f in range(m.n_fronts())]
    P0 = [m.fetch_pivots(f) for f in range(m.n_fronts())]
    m.set_tuning(4, 64)
    assert m.lu_piv_with_free() == 0
    small = 0
    for f in range(m.n_fronts()):
        info = m.front_info(f)
        if info["fs"] and info["fs"] + info["ss"] <= 64:
            small += 1
        assert np.array_equal(P0[f], m.fetch_pivots(f)), f
        G = m.fetch_front(f)
        if G.size:
            err = np.abs(G - F0[f]).max() / max(1e-300, np.abs(F0[f]).max())
            assert err <= (1e-12 if prec == U.PREC_FP64 else 1e-5), (f, err)
    say("  fronts through the whole-front kernel:", small, "of", m.n_fronts())
    assert small > 0
    m.solve_lu_piv()
    m.check_lu_piv()
    if prec == U.PREC_FP64:
        assert m.backward_error() <= ETA_TOL


@pytest.mark.parametrize("prec", [U.PREC_FP64, U.PREC_MIXED])
@pytest.mark.parametrize("group", [1, 4])
def test_cholesky_lookahead_matches_in_order(lib, prec, group):
    """The lookahead schedule (next panel's diagonal block and TRSM on a
    second stream) factors to the in-order schedule's values (rounding only:
    the next panel's column block uses 64-wide tiles)."""
    m = U.Mesh()
    m.build_synthetic(2, (8, 8, 7), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=17, alpha=64.0, scale_kind=1, p0=0.5)
    m.random_matrix()
    m.set_rhs()
    m.set_precision(prec, 1)
    m.set_tuning(1, group)
    m.set_tuning(5, 0)
    assert m.chol_with_free() == 0
    ref = [np.tril(m.fetch_front(f)) for f in range(m.n_fronts())]
    m.set_tuning(5, 1)
    for rep in range(2):
        assert m.chol_with_free() == 0
        worst = 0.0
        for f in range(m.n_fronts()):
            G = np.tril(m.fetch_front(f))
            if G.size:
                worst = max(worst, np.abs(G - ref[f]).max() / max(1e-300, np.abs(ref[f]).max()))
        say("  lookahead vs in order, worst front rel diff %.3e" % worst)
        assert worst <= (1e-12 if prec == U.PREC_FP64 else 1e-5)
    m.solve_chol()
    m.check_chol()
    if prec == U.PREC_FP64:
        assert m.backward_error() <= ETA_TOL


def _chol_mesh(seed, alpha, beta=0.0):
    m = U.Mesh()
    m.build_synthetic(2, (8, 8, 7), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=seed, alpha=alpha, beta=beta, scale_kind=1, p0=0.5)
    m.random_matrix()
    m.set_rhs()
    return m


@pytest.mark.parametrize("prec", [U.PREC_FP64, U.PREC_MIXED])
def test_cholesky_fused_diagonal_same_bits(lib, prec):
    """Diagonal blocks factored by the workgroup of the panel update that
    produced them (no diagonal launch inside a panel group) give the same
    front values, bit for bit, as one diagonal launch per panel, over the
    part of a front the Cholesky path defines: the lower triangle, and W in
    the upper triangles of the 64 x 64 diagonal blocks of ATL (the rest of
    the upper triangle is never assembled or read and holds stale values)."""
    def defined(m, f):
        G = m.fetch_front(f)
        fs = m.front_info(f)["fs"]
        return [np.tril(G)] + [G[k0:min(k0 + 64, fs), k0:min(k0 + 64, fs)].copy() for k0 in range(0, fs, 64)]

    m = _chol_mesh(19, 64.0)
    m.set_precision(prec, 1)
    say("case fused diagonal prec", prec, "max fs", max(m.front_info(f)["fs"] for f in range(m.n_fronts())))
    m.set_tuning(7, 0)  # per-panel launches (the path keys 5 and 6 shape)
    m.set_tuning(6, 0)
    assert m.chol_with_free() == 0
    ref = [defined(m, f) for f in range(m.n_fronts())]
    m.set_tuning(6, 1)
    assert m.get_tuning(6) == 1
    for rep in range(2):
        assert m.chol_with_free() == 0
        for f in range(m.n_fronts()):
            got = defined(m, f)
            assert all(np.array_equal(a, b) for a, b in zip(got, ref[f])), (rep, f)
    m.solve_chol()
    m.check_chol()
    if prec == U.PREC_FP64:
        assert m.backward_error() <= ETA_TOL


def test_cholesky_fused_diagonal_reports_failure(lib):
    """A matrix that is not positive definite fails with the same element
    reported whether or not the diagonal blocks are fused.  Element matrices
    M^T M + 50 S (no shift): the Cholesky path reads their lower triangle,
    i.e. the symmetric matrix whose strict lower part carries 50 S, whose
    eigenvalues (~ +-50 sqrt(2 m)) swamp those of M^T M (<= ~4 m) for m = 64."""
    got = []
    for fuse in (0, 1):
        m = _chol_mesh(23, 0.0, 50.0)
        m.set_tuning(7, 0)
        m.set_tuning(6, fuse)
        got.append(m.chol_with_free())
    m = _chol_mesh(23, 0.0, 50.0)
    m.set_tuning(7, 1)
    got.append(m.chol_with_free())
    say("  info unfused %d fused %d group launches %d" % tuple(got))
    assert got[0] != 0 and got[0] == got[1] == got[2]


@pytest.mark.parametrize("prec", [U.PREC_FP64, U.PREC_MIXED])
@pytest.mark.parametrize("group", [1, 3, 8, 16])
def test_cholesky_group_launch_matches_panel_path(lib, prec, group):
    """Each panel group in one persistent launch (key 7, default; the
    workgroups of one front's row tiles wait for each other's diagonal
    blocks inside the launch, left-looking within the group) against the
    per-panel launches (key 7 = 0, right-looking): the same factors up to
    rounding (the update order inside a group differs), run after run the
    same bits, and the FP64 path's backward error at the FP64 bound."""
    m = _chol_mesh(29, 64.0)
    m.set_precision(prec, 1)
    m.set_tuning(1, group)
    say("case group launch prec", prec, "group", group)
    m.set_tuning(7, 0)
    assert m.chol_with_free() == 0
    nf = m.n_fronts()
    ref = [np.tril(m.fetch_front(f)) for f in range(nf)]
    m.set_tuning(7, 1)  # every level (the default, 2, only the levels with few fronts)
    assert m.chol_with_free() == 0
    got = [np.tril(m.fetch_front(f)) for f in range(nf)]
    tol = 1e-12 if prec == U.PREC_FP64 else 2e-5
    worst = max(float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)) for a, b in zip(got, ref))
    say("  worst front deviation %.2e" % worst)
    assert worst <= tol
    assert m.chol_with_free() == 0
    again = [np.tril(m.fetch_front(f)) for f in range(nf)]
    assert all(np.array_equal(a, b) for a, b in zip(again, got))
    m.solve_chol()
    m.check_chol()
    if prec == U.PREC_FP64:
        assert m.backward_error() <= ETA_TOL
    else:
        m.improve(U.UHM_CHOL, 1e-12, 30)
        assert m.backward_error() <= 1e-12


@pytest.mark.parametrize("method", [U.UHM_LU_PIV, U.UHM_CHOL])
def test_coarsening_tree_factor_solve_vs_oracle(lib, oracle_lib, method):
    """A mesh of leaf elements only, its tree from uhm_build_tree_ (the
    reference's METIS-free coarsening builder), factorized and solved on the
    GPU: backward error <= 1e-13 and the solution equal to the oracle's on
    the same tree and element matrices."""
    from meshes import MeshSpec, qp_spec
    from oracle.oracle import OracleMesh
    spec = qp_spec((9, 7), 3)
    leaves = MeshSpec(spec.nodes, spec.elems, [-1] * len(spec.elems), [[] for _ in spec.elems])
    m, _ = leaves.build_product(U)
    m.build_tree()
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=8, alpha=16.0, beta=0.1 if method == U.UHM_LU_PIV else 0.0, scale_kind=1, p0=0.5)
    m.random_matrix()
    m.set_rhs()
    if method == U.UHM_LU_PIV:
        assert m.lu_piv_with_free() == 0
        m.solve_lu_piv()
        m.check_lu_piv()
    else:
        assert m.chol_with_free() == 0
        m.solve_chol()
        m.check_chol()
    assert m.backward_error() <= 1e-13
    om = OracleMesh.from_connectivity(m.export_connectivity())
    om.init_numeric(1)
    for f in range(m.n_fronts()):
        fi = m.front_info(f)
        if fi["leaf"]:
            om.set_front(fi["id"], m.fetch_stage(f))
            om.set_rhs_front(fi["id"], m.fetch_rhs_front(f))
    assert om.factor(method, 64) == 0
    om.solve(method)
    xg = m.fetch_x()[:, 0]
    xo = om.x_global(len(xg))[:, 0]
    assert np.abs(xg - xo).max() <= 1e-10 * np.abs(xo).max()
