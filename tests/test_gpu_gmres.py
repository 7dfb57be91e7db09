"""GMRES-IR: GMRES on the left-preconditioned system M^-1 A d = M^-1 r with
M the low-precision factors, every right-hand side in lockstep.

Bars: the refined solution reaches normwise backward error <= 1e-12 and
agrees with the FP64 path's solution of the same input to a
condition-aware forward-error bar (max(1e-10, 100 kappa_1(A) 1e-12));
GMRES-IR needs at most two M^-1 applications more than classical refinement
where both converge; a refinement repeated from the same start gives the same
bits; the automatic kind switches to GMRES-IR when classical steps
contract slowly."""
import sys

import numpy as np
import pytest

import uhm_amd as U
from test_gpu_parity import kappa_1

pytestmark = pytest.mark.gpu

ETA = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _varp(nrhs, seed=5):
    """2D variable-order quads (p_e ~ U{1..8}), K_e = M^T M + n I + 0.1 S, LU."""
    m = U.Mesh()
    m.build_synthetic(3, (24, 24), 8, 12345)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, nrhs)
    m.create_matrix_buffer()
    m.set_random_params(seed=seed, alpha=-1.0, beta=0.1, scale_kind=0, p0=0.0, p1=0.0)
    m.random_matrix()
    m.set_rhs()
    return m


def _chol3d(nrhs, seed=6):
    m = U.Mesh()
    m.build_synthetic(2, (6, 5, 5), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, nrhs)
    m.create_matrix_buffer()
    m.set_random_params(seed=seed, alpha=1.0, beta=0.0, scale_kind=1, p0=1.0, p1=0.0)
    m.random_matrix()
    m.set_rhs()
    return m


def _factor(m, method):
    return m.chol_with_free() if method == U.UHM_CHOL else m.lu_piv_with_free()


def _solve(m, method):
    (m.solve_chol if method == U.UHM_CHOL else m.solve_lu_piv)()


CASES = [
    ("varp-lu-8", _varp, 8, U.UHM_LU_PIV),
    ("varp-lu-1", _varp, 1, U.UHM_LU_PIV),
    ("chol3d-12", _chol3d, 12, U.UHM_CHOL),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_gmres_ir_reaches_fp64_accuracy(lib, case):
    name, make, nrhs, method = case
    say("case", name)
    m = make(nrhs)
    assert _factor(m, method) == 0
    _solve(m, method)
    x64 = m.fetch_x().copy()
    kappa = kappa_1(m)
    m.set_precision(U.PREC_MIXED, 0)
    assert _factor(m, method) == 0
    # classical
    _solve(m, method)
    m.set_refinement(U.REFINE_CLASSICAL)
    nc = m.improve(method, ETA, 60)
    eta_c, _ = m.refine_history()
    sc = m.refine_solves()
    say("  classical: %d corrections, %d solves, eta %s" % (nc, sc, " ".join("%.1e" % e for e in eta_c)))
    # GMRES-IR
    _solve(m, method)
    m.set_refinement(U.REFINE_GMRES, 20, 1e-6)
    ng = m.improve(method, ETA, 10)
    eta_g, _ = m.refine_history()
    sg = m.refine_solves()
    say("  GMRES-IR: %d corrections, %d solves, eta %s" % (ng, sg, " ".join("%.1e" % e for e in eta_g)))
    assert eta_g[0] > ETA
    assert m.backward_error() <= ETA
    if eta_c[-1] <= ETA:
        # one outer GMRES-IR step costs 1 + (inner steps) applications of
        # M^-1; on a system that classical refinement already solves in a
        # few steps it may use those few more
        assert sg <= max(sc, 1) + 2
    x = m.fetch_x()
    err = (np.abs(x - x64).max(axis=0) / np.abs(x64).max(axis=0)).max()
    tol = max(1e-10, 100.0 * kappa * ETA)
    say("  kappa_1 ~ %.2e, forward error vs FP64 path %.3e (bar %.2e)" % (kappa, err, tol))
    assert err <= tol
    # copy_out sees the refined solution on every front
    m.check_chol() if method == U.UHM_CHOL else m.check_lu_piv()
    assert m.backward_error() <= ETA


def test_gmres_ir_repeatable(lib):
    method = U.UHM_LU_PIV
    m = _varp(8, seed=7)
    m.set_precision(U.PREC_MIXED, 0)
    m.set_refinement(U.REFINE_GMRES, 10, 1e-5)
    assert _factor(m, method) == 0
    xs = []
    for rep in range(2):
        _solve(m, method)
        m.improve(method, ETA, 10)
        xs.append(m.fetch_x().copy())
    assert np.array_equal(xs[0], xs[1])
    assert m.backward_error() <= ETA


def test_auto_refinement(lib):
    method = U.UHM_LU_PIV
    m = _varp(4, seed=9)
    m.set_precision(U.PREC_MIXED, 0)
    m.set_refinement(U.REFINE_AUTO, 20, 1e-6)
    assert _factor(m, method) == 0
    _solve(m, method)
    n = m.improve(method, ETA, 20)
    eta, _ = m.refine_history()
    say("  auto: %d corrections, %d solves, eta %s" % (n, m.refine_solves(), " ".join("%.1e" % e for e in eta)))
    assert m.backward_error() <= ETA


def _lu3d(nrhs, seed=10):
    m = U.Mesh()
    m.build_synthetic(2, (6, 5, 5), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, nrhs)
    m.create_matrix_buffer()
    m.set_random_params(seed=seed, alpha=64.0, beta=0.1, scale_kind=1, p0=0.5)
    m.random_matrix()
    m.set_rhs()
    return m


@pytest.mark.parametrize("case", [("chol3d-bf16", _chol3d, U.UHM_CHOL), ("lu3d-bf16", _lu3d, U.UHM_LU_PIV)],
                         ids=lambda c: c[0])
@pytest.mark.parametrize("top", [0, 1])
def test_bf16_schur_refined_to_fp64_accuracy(lib, case, top):
    """BF16- and FP16-operand Schur complement products (FP32 accumulation),
    FP32 fronts below the top `top` levels: GMRES-IR brings every right-hand side
    to backward error <= 1e-12 and to the FP64 path's solution within the
    condition-aware bar, and the unrefined BF16 solution is measurably less
    accurate than the FP32-Schur one (the BF16 products really ran)."""
    name, make, method = case
    say("case", name, "top", top)
    nrhs = 6
    m = make(nrhs)
    assert _factor(m, method) == 0
    _solve(m, method)
    x64 = m.fetch_x().copy()
    kappa = kappa_1(m)
    eta0 = {}
    # low-precision products on every level (the default keeps them to
    # levels with large Schur blocks, which meshes this small do not have)
    m.set_tuning(10, 0)
    for mode in (U.PREC_MIXED, U.PREC_MIXED_BF16, U.PREC_MIXED_FP16):
        m.set_precision(mode, top)
        assert _factor(m, method) == 0
        _solve(m, method)
        m.set_refinement(U.REFINE_GMRES, 30, 1e-4)
        n = m.improve(method, ETA, 10)
        eta, _ = m.refine_history()
        eta0[mode] = eta[0]
        say("  mode %d: %d corrections, %d solves, eta %s" % (mode, n, m.refine_solves(),
                                                                " ".join("%.1e" % e for e in eta)))
        assert m.backward_error() <= ETA
        x = m.fetch_x()
        err = (np.abs(x - x64).max(axis=0) / np.abs(x64).max(axis=0)).max()
        tol = max(1e-10, 100.0 * kappa * ETA)
        say("  kappa_1 ~ %.2e, forward error vs FP64 path %.3e (bar %.2e)" % (kappa, err, tol))
        assert err <= tol
    assert eta0[U.PREC_MIXED_BF16] > 10.0 * eta0[U.PREC_MIXED]
    # FP16 keeps three more bits than BF16: its unrefined solution sits between
    assert eta0[U.PREC_MIXED_FP16] > eta0[U.PREC_MIXED] and eta0[U.PREC_MIXED_FP16] < eta0[U.PREC_MIXED_BF16]


def test_low_precision_level_threshold(lib):
    """Tuning key 10: packed low-precision Schur products only on the levels
    whose Schur blocks average at least that many rows (default 768); the
    other levels take FP32 products, so with a threshold above every level
    mode 2 gives mode 1's factors bit for bit, and with 0 every level packs."""
    m = _chol3d(1)
    assert m.get_tuning(10) == 768
    m.set_profile(True)
    m.set_precision(U.PREC_MIXED, 0)
    assert m.chol_with_free() == 0
    ref = [np.tril(m.fetch_front(f)) for f in range(m.n_fronts())]
    m.set_precision(U.PREC_MIXED_BF16, 0)
    m.set_tuning(10, 1 << 30)
    assert m.chol_with_free() == 0
    assert m.kernel_stats()["lp_pack"][3] == 0
    assert all(np.array_equal(np.tril(m.fetch_front(f)), ref[f]) for f in range(m.n_fronts()))
    m.set_tuning(10, 0)
    assert m.chol_with_free() == 0
    assert m.kernel_stats()["lp_pack"][3] > 0
    assert any(not np.array_equal(np.tril(m.fetch_front(f)), ref[f]) for f in range(m.n_fronts()))
