"""Block solves (many right-hand sides as FP64 MFMA products over 64-column
tiles) against the per-column solves of the same factorization and against
the oracle.

Bars: the block solution agrees with the per-column solution of the same
factors to 1e-12 relative per column (both are FP64 arithmetic on the same
factor bits; only the summation order differs) and with the oracle's to
1e-12 relative; normwise backward error <= 1e-13 on FP64 factors; graph
replays of a block solve reproduce the eager bits; mixed-precision
refinement through block solves reaches 1e-12."""
import sys

import numpy as np
import pytest

import uhm_amd as U
from test_gpu_parity import _oracle_from_product

pytestmark = pytest.mark.gpu

PER_COLUMN = 10 ** 6   # block-solve threshold above every n_rhs: per-column solves
BLOCK = 1              # block solves for any n_rhs


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mesh(dims, p, method, nrhs, seed=7):
    m = U.Mesh()
    m.build_synthetic(2, dims, p, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, nrhs)
    m.create_matrix_buffer()
    m.set_random_params(seed=seed, alpha=64.0, beta=0.1 if method == U.UHM_LU_PIV else 0.0, scale_kind=1, p0=0.5)
    m.random_matrix()
    m.set_rhs()
    return m


def _factor(m, method):
    return m.chol_with_free() if method == U.UHM_CHOL else m.lu_piv_with_free()


def _solve(m, method):
    (m.solve_chol if method == U.UHM_CHOL else m.solve_lu_piv)()


def _check(m, method):
    (m.check_chol if method == U.UHM_CHOL else m.check_lu_piv)()


CASES = [
    # name, dims, p, method, mixed top levels (None: FP64), panel width, n_rhs
    ("lu-fp64-70", (6, 5, 5), 3, U.UHM_LU_PIV, None, 64, 70),
    ("chol-w-fp64-70", (6, 5, 5), 3, U.UHM_CHOL, None, 64, 70),
    ("chol-nb32-16", (5, 5, 4), 3, U.UHM_CHOL, None, 32, 16),
    ("lu-nb32-9", (5, 4, 4), 3, U.UHM_LU_PIV, None, 32, 9),
    ("lu-mixed-64", (5, 5, 4), 3, U.UHM_LU_PIV, 1, 64, 64),
    ("chol-mixed-130", (6, 5, 5), 3, U.UHM_CHOL, 1, 64, 130),
    # top fronts of a few hundred rows: the level products split K
    ("chol-w-split-70", (8, 8, 7), 3, U.UHM_CHOL, None, 64, 70),
    ("lu-split-40", (8, 7, 7), 3, U.UHM_LU_PIV, None, 64, 40),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_block_solve_matches_per_column(lib, case):
    name, dims, p, method, top, nb, nrhs = case
    say("case", name)
    nb0 = U.get_hier_block_size()
    U.set_hier_block_size(nb)
    try:
        m = _mesh(dims, p, method, nrhs)
        if top is not None:
            m.set_precision(U.PREC_MIXED, top)
        assert _factor(m, method) == 0
        m.set_tuning(3, PER_COLUMN)
        _solve(m, method)
        xs = m.fetch_x().copy()
        m.set_tuning(3, BLOCK)
        _solve(m, method)
        xb = m.fetch_x().copy()
    finally:
        U.set_hier_block_size(nb0)
    assert xb.shape == (m.get_n_dof(), nrhs)
    err = (np.abs(xb - xs).max(axis=0) / np.abs(xs).max(axis=0)).max()
    say("  max fs", max(m.front_info(f)["fs"] for f in range(m.n_fronts())),
        " worst column rel diff block vs per-column %.3e" % err)
    # FP32 levels: the block path multiplies in FP32 (the factor's own
    # accuracy), the per-column path in FP64 on the same FP32 entries
    assert err <= (1e-12 if top is None else 1e-5)
    assert not np.array_equal(xb[:, 0], xb[:, 1])
    if top is None:
        _check(m, method)
        say("  backward error %.3e" % m.backward_error())
        assert m.backward_error() <= 1e-13


@pytest.mark.parametrize("method", [U.UHM_LU_PIV, U.UHM_CHOL])
def test_block_solve_vs_oracle(lib, oracle_lib, method):
    m = _mesh((5, 4, 4), 3, method, 12, seed=3)
    m.set_tuning(3, BLOCK)
    assert _factor(m, method) == 0
    _solve(m, method)
    xg = m.fetch_x()
    spec, om = _oracle_from_product(m, method)
    for f in range(m.n_fronts()):
        info = m.front_info(f)
        if info["leaf"]:
            om.set_rhs_front(info["id"], m.fetch_rhs_front(f))
    assert om.factor(method, 64) == 0
    om.solve(method)
    xo = om.x_global(m.get_n_dof())
    err = (np.abs(xg - xo).max(axis=0) / np.abs(xo).max(axis=0)).max()
    say("  worst column x rel err vs oracle %.3e" % err)
    assert err <= 1e-12


def test_block_solve_graph_replays(lib):
    method = U.UHM_CHOL
    m = _mesh((5, 5, 4), 3, method, 70, seed=4)
    m.set_tuning(3, BLOCK)
    assert _factor(m, method) == 0
    _solve(m, method)
    x0 = m.fetch_x().copy()
    m.set_graph(True)
    for rep in range(3):
        assert _factor(m, method) == 0
        _solve(m, method)
        assert np.array_equal(m.fetch_x(), x0), rep
        _check(m, method)
        assert m.backward_error() <= 1e-13


@pytest.mark.parametrize("method", [U.UHM_LU_PIV, U.UHM_CHOL])
def test_block_solve_mixed_refinement(lib, method):
    m = _mesh((6, 5, 5), 3, method, 40, seed=8)
    m.set_precision(U.PREC_MIXED, 1)
    m.set_tuning(3, BLOCK)
    assert _factor(m, method) == 0
    _solve(m, method)
    n = m.improve(method, 1e-12, 30)
    eta, ms = m.refine_history()
    say("  corrections", n, " ".join("%.2e" % e for e in eta))
    assert m.backward_error() <= 1e-12
    assert eta[0] > 1e-12 * 1e-3


@pytest.mark.parametrize("nrhs", [1, 70])
def test_leaf_residual_matches_csr(lib, nrhs):
    """The residual from the leaf products (check / refinement) against
    b - A x on the host with the assembled CSR matrix, on a mixed-precision
    solution whose residual is far above rounding (so the two summation
    orders must agree to 1e-6 relative)."""
    method = U.UHM_LU_PIV
    m = _mesh((5, 5, 4), 3, method, nrhs, seed=12)
    m.set_precision(U.PREC_MIXED, 0)
    assert _factor(m, method) == 0
    _solve(m, method)
    _check(m, method)
    eta_dev, r1_dev = m.backward_error(), m.get_residual()
    rp, col, val = m.fetch_csr()
    x, b = m.fetch_x(), m.fetch_b()
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    Ax = np.zeros_like(x)
    for c in range(nrhs):
        Ax[:, c] = np.bincount(rows, weights=val * x[col, c], minlength=n)
    r = b - Ax
    amax = np.bincount(rows, weights=np.abs(val), minlength=n).max()
    eta = max(np.abs(r[:, c]).max() / (amax * np.abs(x[:, c]).max() + np.abs(b[:, c]).max()) for c in range(nrhs))
    r1 = np.abs(r).sum()
    say("  eta device %.6e host %.6e   |r|_1 device %.6e host %.6e" % (eta_dev, eta, r1_dev, r1))
    assert eta > 1e-12
    assert abs(eta_dev - eta) <= 1e-6 * eta
    assert abs(r1_dev - r1) <= 1e-6 * r1


@pytest.mark.parametrize("prec", [U.PREC_FP64, U.PREC_MIXED])
def test_small_front_block_steps_same_bits(lib, prec):
    """The LU block steps of fronts with fs <= 32 by one wavefront per item
    (tuning key 14) give the same solution bits as the general block-step
    kernel, on a variable-order 2D mesh whose lower levels are all small."""
    m = U.Mesh()
    m.build_synthetic(3, (24, 20), 8, 12345)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 70)
    m.create_matrix_buffer()
    m.set_random_params(seed=4, alpha=-1.0, beta=0.1)
    m.random_matrix()
    m.set_rhs()
    m.set_precision(prec, 1)
    assert m.lu_piv_with_free() == 0
    m.set_tuning(14, 0)
    m.solve_lu_piv()
    x0 = m.fetch_x().copy()
    m.set_tuning(14, 1)
    m.solve_lu_piv()
    x1 = m.fetch_x()
    assert np.array_equal(x0.view(np.uint64), x1.view(np.uint64))
    if prec == U.PREC_FP64:
        m.check_lu_piv()
        assert m.backward_error() <= 1e-13
