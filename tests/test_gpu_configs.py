"""Parity at the full sizes of the benchmark configurations.

cfg2  3D 16^3 hex Q3 (117,649 dofs), K_e = (M^T M + 64 I + 0.1 S) lognormal(0,1),
      LU with partial pivoting, FP64 on the GPU: every front's pivot sequence
      equals the oracle's (the C restatement of lu_piv_hier with the
      reference's 256-wide tiles) except where the oracle's own choice was a
      near tie (relative margin below a rounding bound), the solution
      matches the oracle's to 1e-12 relative, backward error <= 1e-13.
cfg3  3D 16^3 hex Q4 (274,625 dofs), SPD K_e = (M^T M + 125 I) 10^U(-2,2),
      Cholesky: FP64 path vs the oracle as for cfg2; the mixed path (FP32
      fronts below the top two levels) refined to backward error <= 1e-12,
      with its forward error against both FP64 solutions.
cfg4  2D 256^2 quads, p_e ~ U{1..8}, K_e = M^T M + n I + 0.1 S, LU, FP32
      fronts, 64 right-hand sides: ten reseeds on the fixed pattern through
      the captured factorization / solve graphs, each refined to <= 1e-12.
cfg5  3D 24^3 hex Q3 (389,017 dofs), SPD K_e = (M^T M + 64 I) lognormal(0,0.5),
      Cholesky with BF16-operand (and FP16-operand) Schur products, FP64 root
      front, 256 right-hand sides, GMRES-IR to <= 1e-12.
"""
import sys
import time

import numpy as np
import pytest

import uhm_amd as U
from oracle.oracle import OracleMesh, set_threads

pytestmark = pytest.mark.gpu

# a pivot choice whose two best candidates differ by less than this
# relative margin may fall either way under a different (equally valid)
# rounding order of the updates before it
PIVOT_TIE = 1e-8


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mesh(kind, dims, p, nrhs, seed, alpha, beta, sk, p0, p1=0.0, mesh_seed=0):
    m = U.Mesh()
    m.build_synthetic(kind, dims, p, mesh_seed)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, nrhs)
    m.create_matrix_buffer()
    m.set_random_params(seed=seed, alpha=alpha, beta=beta, scale_kind=sk, p0=p0, p1=p1)
    m.random_matrix()
    m.set_rhs()
    return m


def _oracle_solution(m, method):
    """The oracle's factorization and solve of the GPU mesh's own inputs
    (leaf stages and right-hand sides fetched from the device)."""
    set_threads(16)
    t0 = time.perf_counter()
    om = OracleMesh.from_connectivity(m.export_connectivity())
    om.init_numeric(m.n_rhs)
    for f in range(m.n_fronts()):
        fi = m.front_info(f)
        if fi["leaf"]:
            om.set_front(fi["id"], m.fetch_stage(f))
            om.set_rhs_front(fi["id"], m.fetch_rhs_front(f))
    t1 = time.perf_counter()
    info = om.factor(method, 256)
    t2 = time.perf_counter()
    om.solve(method)
    say("  oracle: inputs %.1f s, factor %.1f s (%.1f GFLOP), solve %.1f s, info %d"
        % (t1 - t0, t2 - t1, om.flops(method) / 1e9, time.perf_counter() - t2, info))
    assert info == 0
    return om, om.x_global(m.get_n_dof())


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_cfg2_lu_fp64_matches_oracle(lib, oracle_lib):
    say("case cfg2: 16^3 Q3 LU FP64")
    m = _mesh(2, (16, 16, 16), 3, 1, seed=2, alpha=64.0, beta=0.1, sk=1, p0=1.0)
    assert m.get_n_dof() == 117649
    assert m.lu_piv_with_free() == 0
    m.solve_lu_piv()
    m.check_lu_piv()
    eta = m.backward_error()
    x = m.fetch_x()
    om, xo = _oracle_solution(m, U.UHM_LU_PIV)
    ties, differ = [], 0
    for f in range(m.n_fronts()):
        e = m.front_info(f)["id"]
        gp, op = m.fetch_pivots(f), om.pivots(e)
        if np.array_equal(gp, op):
            continue
        differ += 1
        k = int(np.nonzero(gp != op)[0][0])
        ties.append((f, k, float(om.margins(e)[k])))
    err = _rel(x, xo)
    say("  eta %.2e, x vs oracle %.2e, fronts with other pivots %d %s" % (eta, err, differ, ties[:8]))
    assert eta <= 1e-13
    assert all(mg < PIVOT_TIE for _, _, mg in ties), ties
    assert err <= 1e-12


def test_cfg3_chol_fp64_and_mixed(lib, oracle_lib):
    say("case cfg3: 16^3 Q4 Cholesky, FP64 and mixed")
    m = _mesh(2, (16, 16, 16), 4, 1, seed=3, alpha=125.0, beta=0.0, sk=2, p0=-2.0, p1=2.0)
    assert m.get_n_dof() == 274625
    assert m.chol_with_free() == 0
    m.solve_chol()
    m.check_chol()
    eta64 = m.backward_error()
    x64 = m.fetch_x()
    om, xo = _oracle_solution(m, U.UHM_CHOL)
    err64 = _rel(x64, xo)
    del om
    m.set_precision(U.PREC_MIXED, 2)
    assert m.chol_with_free() == 0
    m.solve_chol()
    n = m.improve(U.UHM_CHOL, 1e-12, 30)
    etam = m.backward_error()
    hist, _ = m.refine_history()
    xm = m.fetch_x()
    fw64, fwo = _rel(xm, x64), _rel(xm, xo)
    say("  fp64: eta %.2e, x vs oracle %.2e; mixed: %d corrections, eta %s, forward error vs fp64 %.2e, "
        "vs oracle %.2e" % (eta64, err64, n, " -> ".join("%.1e" % h for h in hist), fw64, fwo))
    assert eta64 <= 1e-13
    assert err64 <= 1e-12
    assert etam <= 1e-12
    assert fw64 <= 1e-12 and fwo <= 1e-12


def test_cfg4_varp_64rhs_graph_reseeds(lib):
    say("case cfg4: 256^2 var-p LU, 64 RHS, 10 graph-replayed reseeds")
    m = _mesh(3, (256, 256), 8, 64, seed=1, alpha=-1.0, beta=0.1, sk=0, p0=0.0, mesh_seed=12345)
    m.set_precision(U.PREC_MIXED, 0)
    m.set_refinement(U.REFINE_AUTO, 20, 1e-6)
    m.set_graph(True)
    for seed in range(1, 11):
        m.set_random_params(seed=seed, alpha=-1.0, beta=0.1, scale_kind=0, p0=0.0, p1=0.0)
        m.random_matrix()
        m.set_rhs()
        assert m.lu_piv_with_free() == 0, seed
        m.solve_lu_piv()
        n = m.improve(U.UHM_LU_PIV, 1e-12, 30)
        eta = m.backward_error()
        say("  reseed %d: %d corrections, %d solves, eta %.2e" % (seed, n, m.refine_solves(), eta))
        assert eta <= 1e-12, seed
    m.set_graph(False)


@pytest.mark.parametrize("mode", [U.PREC_MIXED_BF16, U.PREC_MIXED_FP16])
def test_cfg5_q3_24cube_low_precision_256rhs(lib, mode):
    say("case cfg5: 24^3 Q3 Cholesky, 256 RHS, precision mode", mode)
    m = _mesh(2, (24, 24, 24), 3, 256, seed=5, alpha=64.0, beta=0.0, sk=1, p0=0.5)
    assert m.get_n_dof() == 389017
    m.set_precision(mode, 1)
    m.set_refinement(U.REFINE_GMRES, 20, 1e-6)
    assert m.chol_with_free() == 0
    m.solve_chol()
    n = m.improve(U.UHM_CHOL, 1e-12, 30)
    hist, _ = m.refine_history()
    eta = m.backward_error()
    say("  %d corrections, %d solves, eta %s" % (n, m.refine_solves(), " -> ".join("%.1e" % h for h in hist)))
    assert eta <= 1e-12
