"""The synthetic element generator (uhmk_generate_elements behind
uhm_random_matrix_).

* Its Philox-4x32-10 restatement (tests/rng_restated.py) reproduces the
  published known-answer vectors of Random123 (kat_vectors, philox4x32_10:
  all-zero, all-ones and pi-digit counters and keys).
* On the GPU: the right-hand sides the device writes are the restated
  normal draws (to the last bits of libm's log / sin / cos), and every leaf
  stage equals K_e = s_e (M^T M + shift I + beta S) rebuilt on the host from
  the restated draws and scattered through the oracle's copy-in map, for
  every scale kind, with and without S, for a positive shift and for
  alpha < 0 (shift = |alpha| x element size).
"""
import numpy as np
import pytest

import uhm_amd as U
from oracle.oracle import OracleMesh
from rng_restated import element_matrix, element_rhs, normals, philox4x32_10

# Random123 kat_vectors (philox4x32 10): counter, key -> output
KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    got = tuple(int(v) for v in philox4x32_10(ctr, key))
    assert got == want, ["%08x" % v for v in got]


def test_normals_distribution():
    # 2^16 draws of one stream: mean, variance and a tail fraction of N(0, 1)
    z = normals(12345, 7, 0, np.arange(1 << 16))
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1.0) < 0.02
    assert abs((np.abs(z) > 1.96).mean() - 0.05) < 0.005


CASES = [
    # kind, dims, p, seed, alpha, beta, scale kind, p0, p1, nrhs
    ("3d-q3-lognormal-S", 2, (3, 2, 2), 3, 41, 64.0, 0.1, 1, 1.0, 0.0, 2),
    ("3d-q2-pow10", 2, (2, 2, 3), 2, 42, 27.0, 0.0, 2, -2.0, 2.0, 1),
    ("2d-varp-negative-alpha", 3, (6, 5), 8, 43, -1.0, 0.1, 0, 0.0, 0.0, 3),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_generated_elements_match_restatement(lib, oracle_lib, case):
    name, kind, dims, p, seed, alpha, beta, sk, p0, p1, nrhs = case
    print("case generator", name, flush=True)
    m = U.Mesh()
    m.build_synthetic(kind, dims, p, 12345 if kind == 3 else 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, nrhs)
    m.create_matrix_buffer()
    m.set_random_params(seed=seed, alpha=alpha, beta=beta, scale_kind=sk, p0=p0, p1=p1)
    m.random_matrix()
    c = m.export_connectivity()
    om = OracleMesh.from_connectivity(c)
    om.init_numeric(nrhs)
    eptr, enode, ndof = c["eptr"], c["enode"], c["ndof"]
    n_leaf = 0
    for f in range(m.n_fronts()):
        fi = m.front_info(f)
        if not fi["leaf"]:
            continue
        e = fi["id"]
        nl = enode[eptr[e]:eptr[e + 1]]
        me = int(ndof[nl].sum())
        assert me == fi["fs"] + fi["ss"]
        K = element_matrix(seed, e, me, alpha, beta, sk, p0, p1)
        om.copy_in(e, nl, U.UHM_LHS, K)
        want = om.input(e)
        got = m.fetch_stage(f)
        # a generator that added alpha itself for alpha < 0 (shift -1
        # instead of me) would be off by (me + 1) s_e on the diagonal
        assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max(), (name, f)
        # the copy-in map (element dof i -> front row), read from the
        # oracle's input of a fresh mesh given diag(1 .. me)
        ob = OracleMesh.from_connectivity(c)
        ob.init_numeric(1)
        ob.copy_in(e, nl, U.UHM_LHS, np.diag(np.arange(1.0, me + 1.0)))
        d = np.diag(ob.input(e))
        front_row = np.empty(me, np.int64)
        front_row[np.rint(d).astype(np.int64) - 1] = np.arange(me)
        b = element_rhs(seed, e, me, nrhs)
        gotb = m.fetch_rhs_front(f)
        assert np.abs(gotb[front_row, :] - b).max() <= 1e-14 * np.abs(b).max(), (name, f)
        n_leaf += 1
    assert n_leaf == sum(1 for f in range(m.n_fronts()) if m.front_info(f)["leaf"]) > 0
