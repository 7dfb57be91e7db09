"""Graph mode: the captured factorization and solves, replayed several times
with eager work and host synchronisation in between, reproduce the eager
results bit for bit, and pick up new element values on the same pattern
(re-factorization on a fixed sparsity pattern)."""
import sys

import numpy as np
import pytest

import uhm_amd as U

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mesh(method, seed, nrhs=1):
    m = U.Mesh()
    m.build_synthetic(2, (4, 4, 4), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, nrhs)
    m.create_matrix_buffer()
    m.set_random_params(seed=seed, alpha=64.0, beta=0.1 if method == U.UHM_LU_PIV else 0.0, scale_kind=1, p0=0.5)
    m.random_matrix()
    m.set_rhs()
    return m


def _run(m, method):
    """Factor + solve; the fronts' defined part (Cholesky: the lower
    triangle, the only part the Cholesky path reads) and the solution."""
    chol = method == U.UHM_CHOL
    assert (m.chol_with_free() if chol else m.lu_piv_with_free()) == 0
    (m.solve_chol if chol else m.solve_lu_piv)()
    fronts = [m.fetch_front(f) for f in range(m.n_fronts())]
    if chol:
        fronts = [np.tril(F) for F in fronts]
    return fronts, m.fetch_x().copy()


@pytest.mark.parametrize("method", [U.UHM_CHOL, U.UHM_LU_PIV])
def test_graph_replays_match_eager(lib, method):
    say("case graph method", method)
    m = _mesh(method, 5)
    F0, x0 = _run(m, method)
    m.set_graph(True)
    for rep in range(3):
        F1, x1 = _run(m, method)
        for f in range(m.n_fronts()):
            assert np.array_equal(F0[f], F1[f]), (rep, f)
        assert np.array_equal(x0, x1), rep
        # eager work between replays
        (m.check_chol if method == U.UHM_CHOL else m.check_lu_piv)()
        assert m.backward_error() <= 1e-13
    # new values on the same pattern: eager and replayed agree
    m.set_random_params(seed=9, alpha=64.0, beta=0.1 if method == U.UHM_LU_PIV else 0.0, scale_kind=1, p0=0.5)
    m.random_matrix()
    m.set_graph(False)
    F2, x2 = _run(m, method)
    assert not np.array_equal(F2[-1], F0[-1])
    m.set_graph(True)
    for rep in range(2):
        F3, x3 = _run(m, method)
        for f in range(m.n_fronts()):
            assert np.array_equal(F2[f], F3[f]), (rep, f)
        assert np.array_equal(x2, x3)


@pytest.mark.parametrize("method", [U.UHM_CHOL, U.UHM_LU_PIV])
def test_graph_replays_new_values_without_recapture(lib, method):
    """Re-factorization on a fixed pattern: graph mode stays on while the
    element values are redrawn, so the factorization and solve graphs
    captured on the first seed are replayed on the new values.  Every reseed
    matches, bit for bit, an eager run of a second mesh built with the same
    seed, and its refined solution reaches backward error 1e-12."""
    chol = method == U.UHM_CHOL
    beta = 0.0 if chol else 0.1
    m = _mesh(method, 5, nrhs=2)
    m.set_precision(U.PREC_MIXED, 1)
    m.set_graph(True)
    _run(m, method)                      # captures the graphs on seed 5
    for seed in (9, 10, 11):
        say("case graph reseed method", method, "seed", seed)
        m.set_random_params(seed=seed, alpha=64.0, beta=beta, scale_kind=1, p0=0.5)
        m.random_matrix()
        m.set_rhs()
        F1, x1 = _run(m, method)         # replayed graphs
        ref = _mesh(method, seed, nrhs=2)
        ref.set_precision(U.PREC_MIXED, 1)
        F0, x0 = _run(ref, method)       # eager
        for f in range(m.n_fronts()):
            assert np.array_equal(F0[f], F1[f]), (seed, f)
        assert np.array_equal(x0, x1), seed
        m.improve(method, 1e-12, 30)
        assert m.backward_error() <= 1e-12, (seed, m.backward_error())


def test_graph_mixed_refinement(lib):
    m = _mesh(U.UHM_CHOL, 6)
    m.set_precision(U.PREC_MIXED, 1)
    F0, x0 = _run(m, U.UHM_CHOL)
    n0 = m.improve(U.UHM_CHOL, 1e-12, 30)
    xr0 = m.fetch_x().copy()
    m.set_graph(True)
    for rep in range(2):
        F1, x1 = _run(m, U.UHM_CHOL)
        for f in range(m.n_fronts()):
            assert np.array_equal(F0[f], F1[f])
        assert np.array_equal(x0, x1)
        assert m.improve(U.UHM_CHOL, 1e-12, 30) == n0
        assert np.array_equal(m.fetch_x(), xr0)
        assert m.backward_error() <= 1e-12
