"""Mixed precision on the GPU: FP32 fronts below the top tree levels, FP32
MFMA Schur updates, FP64 iterative refinement against the assembled matrix.

Bars: the low-precision factors agree with the FP64 factorization of the
same input to FP32 accuracy (relative 1e-4 per front, which pins the FP32
storage paths of assembly, pivot blocks, TRSM and GEMM); refinement reaches
normwise backward error <= 1e-12; the refined solution agrees with the FP64
path's solution of the same input to a condition-aware forward-error bar
(max(1e-10, 100 kappa_1(A) 1e-12))."""
import sys

import numpy as np
import pytest

import uhm_amd as U
from test_gpu_parity import kappa_1

pytestmark = pytest.mark.gpu

MIXED_FACTOR_TOL = 1e-4
ETA_TARGET = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mesh(dims, p, method, seed, alpha, beta, sk, p0, p1, nrhs=1):
    m = U.Mesh()
    m.build_synthetic(2, dims, p, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, nrhs)
    m.create_matrix_buffer()
    m.set_random_params(seed=seed, alpha=alpha, beta=beta, scale_kind=sk, p0=p0, p1=p1)
    m.random_matrix()
    m.set_rhs()
    return m


def _factor(m, method):
    return m.chol_with_free() if method == U.UHM_CHOL else m.lu_piv_with_free()


def _solve(m, method):
    (m.solve_chol if method == U.UHM_CHOL else m.solve_lu_piv)()


CASES = [
    # name, dims, p, method, alpha, beta, scale kind, p0, p1, top FP64 levels
    ("chol-q4-top2", (4, 4, 4), 4, U.UHM_CHOL, 125.0, 0.0, 2, -2.0, 2.0, 2),
    ("chol-q3-all32", (6, 6, 5), 3, U.UHM_CHOL, 64.0, 0.0, 1, 0.5, 0.0, 0),
    ("lu-q3-top2", (5, 4, 4), 3, U.UHM_LU_PIV, 64.0, 0.1, 1, 1.0, 0.0, 2),
    ("lu-q3-all32", (6, 5, 5), 3, U.UHM_LU_PIV, 64.0, 0.1, 1, 1.0, 0.0, 0),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_mixed_refines_to_fp64_solution(lib, case):
    name, dims, p, method, alpha, beta, sk, p0, p1, top = case
    say("case", name)
    m = _mesh(dims, p, method, 21, alpha, beta, sk, p0, p1)
    # FP64 reference on the same input
    assert _factor(m, method) == 0
    F64 = [m.fetch_front(f) for f in range(m.n_fronts())]
    _solve(m, method)
    x64 = m.fetch_x()[:, 0].copy()
    # mixed
    m.set_precision(U.PREC_MIXED, top)
    assert _factor(m, method) == 0
    worst = 0.0
    levels = [m.front_info(f)["level"] for f in range(m.n_fronts())]
    top_level = max(levels)
    n32 = 0
    for f in range(m.n_fronts()):
        G = m.fetch_front(f)
        R = F64[f]
        if G.size == 0:
            continue
        if method == U.UHM_CHOL:
            lo = np.tril(np.ones(G.shape, bool))
            G, R = G[lo], R[lo]
        if levels[f] <= top_level - top:
            n32 += 1
            # FP32 storage: every entry is an FP32 number
            assert np.array_equal(G, G.astype(np.float32).astype(np.float64))
        err = np.abs(G - R).max() / max(1e-300, np.abs(R).max())
        worst = max(worst, err)
    say("  FP32 fronts", n32, "of", m.n_fronts(), " worst front rel diff vs FP64", worst)
    assert n32 > 0
    assert worst <= MIXED_FACTOR_TOL
    _solve(m, method)
    n = m.improve(method, ETA_TARGET, 30)
    eta, ms = m.refine_history()
    say("  corrections", n, "backward errors", " ".join("%.2e" % e for e in eta), "(%.2f ms)" % ms)
    assert eta[0] > ETA_TARGET * 1e-3  # the unrefined mixed solution is visibly less accurate
    assert m.backward_error() <= ETA_TARGET
    assert 1 <= n <= 30
    x = m.fetch_x()[:, 0]
    err = np.abs(x - x64).max() / np.abs(x64).max()
    kappa = kappa_1(m)
    tol = max(1e-10, 100.0 * kappa * ETA_TARGET)
    say("  kappa_1 ~ %.2e, forward error vs FP64 path %.3e (bar %.2e)" % (kappa, err, tol))
    assert err <= tol
    # copy_out sees the refined solution on every front (branch pass)
    m.check_chol() if method == U.UHM_CHOL else m.check_lu_piv()
    assert m.backward_error() <= ETA_TARGET


def test_refine_on_fp64_factors_is_a_no_op(lib):
    m = _mesh((4, 4, 4), 3, U.UHM_LU_PIV, 3, 64.0, 0.1, 1, 1.0, 0.0)
    assert m.lu_piv_with_free() == 0
    m.solve_lu_piv()
    x = m.fetch_x()[:, 0].copy()
    assert m.improve(U.UHM_LU_PIV, ETA_TARGET, 10) == 0
    eta, _ = m.refine_history()
    assert len(eta) == 1 and eta[0] <= 1e-14
    assert np.array_equal(m.fetch_x()[:, 0], x)


def test_mixed_multiple_rhs(lib):
    m = _mesh((5, 4, 4), 3, U.UHM_CHOL, 8, 64.0, 0.0, 1, 0.5, 0.0, nrhs=3)
    m.set_precision(U.PREC_MIXED, 1)
    assert m.chol_with_free() == 0
    m.solve_chol()
    n = m.improve(U.UHM_CHOL, ETA_TARGET, 30)
    eta, _ = m.refine_history()
    say("  corrections", n, " ".join("%.2e" % e for e in eta))
    assert m.backward_error() <= ETA_TARGET
    m.check_chol()
    assert m.backward_error() <= ETA_TARGET


def test_precision_switch_back_to_fp64(lib):
    """FP64 -> mixed -> FP64 on one mesh gives the FP64 bits again."""
    m = _mesh((4, 4, 3), 3, U.UHM_LU_PIV, 4, 64.0, 0.1, 1, 1.0, 0.0)
    assert m.lu_piv_with_free() == 0
    F1 = [m.fetch_front(f) for f in range(m.n_fronts())]
    m.set_precision(U.PREC_MIXED, 1)
    assert m.lu_piv_with_free() == 0
    m.set_precision(U.PREC_FP64, 1)
    assert m.lu_piv_with_free() == 0
    for f in range(m.n_fronts()):
        assert np.array_equal(F1[f], m.fetch_front(f))
