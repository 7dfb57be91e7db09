"""Cholesky solves with few right-hand sides: the one-launch-per-level
kernels (tuning key 8 = 1, the default; in-launch dependencies between the
64-row blocks of a front) against the launches per 64-row block step (key 8
= 0) on the same factors.

The two forms sum the block products in different orders, so x agrees to
rounding: 1e-12 relative (normwise, per column), on fronts whose ABL holds
many 64-row chunks (the backward level kernel reads those straight from
memory) and with the top levels in FP64 or mixed storage; the level
solution's normwise backward error is <= 1e-13 on FP64 factors."""
import sys

import numpy as np
import pytest

import uhm_amd as U

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("prec", [U.PREC_FP64, U.PREC_MIXED])
@pytest.mark.parametrize("nrhs", [1, 3])
def test_level_solves_match_block_steps(lib, prec, nrhs):
    m = U.Mesh()
    m.build_synthetic(2, (10, 10, 9), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, nrhs)
    m.create_matrix_buffer()
    m.set_random_params(seed=11, alpha=64.0, beta=0.0, scale_kind=1, p0=0.5)
    m.random_matrix()
    m.set_rhs()
    m.set_precision(prec, 2)
    assert m.chol_with_free() == 0
    big = max(m.front_info(f)["ss"] for f in range(m.n_fronts()))
    say("  largest ABL: %d rows" % big)
    assert big > 5 * 64
    m.set_tuning(8, 0)
    m.solve_chol()
    x0 = m.fetch_x().copy()
    m.set_tuning(8, 1)
    m.solve_chol()
    x1 = m.fetch_x().copy()
    for c in range(nrhs):
        err = np.abs(x1[:, c] - x0[:, c]).max() / np.abs(x0[:, c]).max()
        say("  rhs %d: level vs block steps %.2e" % (c, err))
        assert err <= 1e-12
    m.check_chol()
    if prec == U.PREC_FP64:
        assert m.backward_error() <= 1e-13
