"""Pin the CPU oracle before trusting it (CPU only).

Each oracle function is checked against something it does not share code
with: golden vectors from LAPACK / numpy (tests/golden), the hand-derived
ex1 maps, an independent LCA formulation of the lock rules, brute-force
dense solves of the assembled system, closed-form flop counts, and algebraic
identities chosen so a dropped term, wrong sign or transposed operand fails.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sla

from meshes import GOLDEN, ex1_spec, lca_lock, load_q2, qp_spec
from oracle import oracle as O


def _fronts(name):
    z = np.load(os.path.join(GOLDEN, name))
    off = 0
    poff = 0
    for n, fs in z["sizes"]:
        fin = z["fin"][off:off + n * n].reshape(n, n, order="F")
        fout = z["fout"][off:off + n * n].reshape(n, n, order="F")
        off += n * n
        piv = None
        if "piv" in z:
            piv = z["piv"][poff:poff + fs]
            poff += fs
        yield int(n), int(fs), fin, fout, piv


@pytest.mark.parametrize("b", [1, 3, 64, 256])
def test_oracle_lu_matches_lapack_golden(oracle_lib, b):
    for n, fs, fin, fout, piv in _fronts("lu_fronts.npz"):
        F, p, info = O.dense_lu(fin, fs, b)
        assert info == 0
        np.testing.assert_array_equal(p, piv)  # pivot sequence bit-exact
        scale = max(1.0, np.abs(fout).max())
        np.testing.assert_allclose(F, fout, rtol=0, atol=1e-9 * scale)


@pytest.mark.parametrize("b", [1, 5, 64, 256])
def test_oracle_chol_matches_numpy_golden(oracle_lib, b):
    for n, fs, fin, fout, _ in _fronts("chol_fronts.npz"):
        F, info = O.dense_chol(fin, fs, b)
        assert info == 0
        lo = np.tril(np.ones((n, n), bool))
        scale = max(1.0, np.abs(fout).max())
        np.testing.assert_allclose(F[lo], fout[lo], rtol=0, atol=1e-9 * scale)


def test_oracle_lu_identities(oracle_lib):
    """P ATL = L U and the Schur complement ABR - ABL ATL^-1 ATR, for a
    front whose pivots are forced away from the diagonal."""
    rng = np.random.default_rng(3)
    n, fs = 90, 70
    A = rng.standard_normal((n, n))
    A[:fs, :fs] += np.diag(np.linspace(0.0, -1e-3, fs))  # keep pivoting active
    F, p, info = O.dense_lu(A, fs, 16)
    L = np.tril(F[:fs, :fs], -1) + np.eye(fs)
    U = np.triu(F[:fs, :fs])
    PA = A[:fs, :].copy()
    for k in range(fs):
        PA[[k, k + p[k]]] = PA[[k + p[k], k]]
    np.testing.assert_allclose(L @ U, PA[:, :fs], atol=1e-11)
    np.testing.assert_allclose(L @ F[:fs, fs:], PA[:, fs:], atol=1e-11)   # ATR = L^-1 P ATR
    np.testing.assert_allclose(F[fs:, :fs] @ U, A[fs:, :fs], atol=1e-11)  # ABL = ABL U^-1
    S = A[fs:, fs:] - A[fs:, :fs] @ np.linalg.solve(A[:fs, :fs], A[:fs, fs:])
    np.testing.assert_allclose(F[fs:, fs:], S, atol=1e-9)
    assert np.any(p != 0)


def test_oracle_zero_pivot_reported(oracle_lib):
    A = np.eye(6)
    A[:, 2] = 0.0  # singular ATL: column 2 vanishes
    _, _, info = O.dense_lu(A, 6, 4)
    assert info == 3
    S = np.eye(5)
    S[3, 3] = -1.0
    _, info = O.dense_chol(S, 5, 2)
    assert info == 4


def _check_lock(spec, om):
    res, off, ndof = lca_lock(spec)
    for e in range(spec.n_elem):
        el = om.element(e)
        r = res[e]
        assert el["generation"] == r["generation"]
        assert el["fs"] == r["fs"] and el["ss"] == r["ss"], e
        assert [spec.nodes[k][0] for k in el["factor"]] == r["factor"]
        assert list(el["factor_off"]) == r["factor_off"]
        assert [spec.nodes[k][0] for k in el["schur"]] == r["schur"]
        assert list(el["schur_off"]) == r["schur_off"]
    for nid, o in off.items():
        assert om.node_offset(spec.seq[nid]) == o
    return res, ndof


def test_oracle_lock_ex1_hand_derived(oracle_lib):
    spec, g = ex1_spec()
    om = spec.oracle()
    for e in range(3):
        el = om.element(e)
        assert el["fs"] == g["fs"][e] and el["ss"] == g["ss"][e] and el["generation"] == g["generation"][e]
        assert [[spec.nodes[k][0], int(o)] for k, o in zip(el["factor"], el["factor_off"])] == g["factor"][e]
        assert [[spec.nodes[k][0], int(o)] for k, o in zip(el["schur"], el["schur_off"])] == g["schur"][e]
        assert om.mapper(e).tolist() == g["mapper"][e]
    for nid, o in g["offset"].items():
        assert om.node_offset(spec.seq[int(nid)]) == o
    _check_lock(spec, om)


@pytest.mark.parametrize("dims,p", [((4, 4), 2), ((5, 3), 1), ((3, 2, 2), 2), ((16, 16), 2), ((4, 4, 4), 3)])
def test_oracle_lock_matches_lca(oracle_lib, dims, p):
    spec = qp_spec(dims, p)
    om = spec.oracle()
    res, ndof = _check_lock(spec, om)
    assert ndof == int(np.prod([d * p + 1 for d in dims]))
    # invariants: every dof factored once; factor offsets precede schur
    # offsets inside every element (the child's lower triangle stays lower)
    assert sum(r["fs"] for r in res) == ndof
    for e in range(spec.n_elem):
        el = om.element(e)
        fo = [om.node_offset(k) for k in el["factor"]]
        so = [om.node_offset(k) for k in el["schur"]]
        if fo and so:
            assert max(fo) < min(so)


def test_oracle_mapper_runs_partition_and_land(oracle_lib):
    spec = qp_spec((6, 5), 2)
    om = spec.oracle()
    for c in range(spec.n_elem):
        p = spec.parent[c]
        if p < 0:
            continue
        ec, ep = om.element(c), om.element(p)
        runs = om.mapper(c)
        # runs tile [0, ss_c) in order
        assert runs[:, 0].tolist() == list(np.cumsum([0] + runs[:-1, 3].tolist()))
        assert runs[:, 3].sum() == ec["ss"]
        # each child schur node lands on its own place in the parent
        pos = {}
        for k, o in zip(ep["factor"], ep["factor_off"]):
            pos[k] = (0, o)
        for k, o in zip(ep["schur"], ep["schur_off"]):
            pos[k] = (1, o)
        flat = []
        for oc, op, fsp, nd in runs:
            flat += [(fsp, op + d) for d in range(nd)]
        want = []
        for k in ec["schur"]:
            fsp, o = pos[k]
            want += [(fsp, o + d) for d in range(spec.ndof[spec.nodes[k][0]])]
        assert flat == want
        # condensed: adjacent runs are not mergeable
        for a, b in zip(runs[:-1], runs[1:]):
            assert a[2] != b[2] or b[1] != a[1] + a[3]


@pytest.mark.parametrize("N,method", [(4, 3), (16, 1)])
def test_oracle_solution_matches_dense_solve(oracle_lib, N, method):
    z = load_q2()
    spec = qp_spec((N, N), 2)
    om = spec.oracle()
    om.init_numeric(1)
    el, K, b = z["N%d_elems" % N], z["N%d_K" % N], z["N%d_b" % N]
    assert el.tolist() == spec.elems
    for e in range(len(el)):
        om.copy_in(e, [spec.seq[n] for n in el[e]], 1, K[e])
        om.copy_in(e, [spec.seq[n] for n in el[e]], 2, b[e].reshape(-1, 1))
    assert om.factor(method, 3) == 0
    om.solve(method)
    x = om.x_global(len(spec.nodes))[:, 0]
    xn = np.zeros(len(spec.nodes))
    for nid, _ in spec.nodes:
        xn[nid] = x[om.node_offset(spec.seq[nid])]
    ref = z["N%d_x" % N]
    assert np.abs(xn - ref).max() / np.abs(ref).max() < 1e-12


def test_oracle_flops_closed_form(oracle_lib):
    # LAWN-41 closed forms summed over an independently computed lock of 3D Q3 on 4^3
    spec = qp_spec((4, 4, 4), 3)
    om = spec.oracle()
    res, _, _ = lca_lock(spec)
    lu = chol = 0.0
    for r in res:
        f, s = float(r["fs"]), float(r["ss"])
        lu += 2 * f ** 3 / 3 - f ** 2 / 2 + 5 * f / 6 + 2 * s * f * f + 2 * s * s * f
        chol += f ** 3 / 3 + f ** 2 / 2 + f / 6 + s * f * f + f * s * (s + 1)
    assert abs(om.flops(3) - lu) <= 1e-9 * lu
    assert abs(om.flops(1) - chol) <= 1e-9 * chol
