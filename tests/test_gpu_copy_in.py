"""Element front-end on the device (copy_in, interf.cxx:108-204 ->
Helper_::merge_A / merge_rhs_b, helper.hxx:191-331): the element blocks of
the copy_in calls are scattered into the leaf stages by one launch
(uhmk_copy_in).  Every stage must equal, bit for bit, the oracle's
restatement of the host scatter fed with the same calls: two calls per
element (the second accumulates, in call order, into the first), issued
interleaved over the elements; right-hand sides accumulate the same way;
after random_matrix (which replaces the stages) the next copy_in starts
from zero."""
import numpy as np
import pytest

import uhm_amd as U
from meshes import load_q2, qp_spec

pytestmark = pytest.mark.gpu


def test_copy_in_scatter_matches_oracle(lib, oracle_lib):
    z = load_q2()
    N = 4
    spec = qp_spec((N, N), 2)
    m, E = spec.build_product(U)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    el, K, b = z["N%d_elems" % N], z["N%d_K" % N], z["N%d_b" % N]
    ne = len(el)
    rng = np.random.default_rng(3)
    K2 = [0.5 * K[e] + rng.standard_normal(K[e].shape) for e in range(ne)]
    om = spec.oracle()
    om.init_numeric(1)
    # first pass forward, second pass backward: the per-element order of the
    # calls is K then K2 for every element
    for e in range(ne):
        m.copy_in(E[e], el[e], U.UHM_LHS, K[e])
        om.copy_in(e, [spec.seq[n] for n in el[e]], 1, K[e])
        m.copy_in(E[e], el[e], U.UHM_RHS, b[e].reshape(-1, 1))
    for e in reversed(range(ne)):
        m.copy_in(E[e], el[e], U.UHM_LHS, K2[e])
        om.copy_in(e, [spec.seq[n] for n in el[e]], 1, K2[e])
    leaves = [f for f in range(m.n_fronts()) if m.front_info(f)["leaf"]]
    assert len(leaves) == ne
    R1 = {}
    for f in leaves:
        e = m.front_info(f)["id"]
        assert np.array_equal(m.fetch_stage(f), om.input(e)), f
        R1[f] = m.fetch_rhs_front(f)
        assert np.abs(R1[f]).max() > 0
    # right-hand sides accumulate: b, then 2 b
    for e in range(ne):
        m.copy_in(E[e], el[e], U.UHM_RHS, 2.0 * b[e].reshape(-1, 1))
    for f in leaves:
        assert np.array_equal(m.fetch_rhs_front(f), R1[f] + 2.0 * R1[f]), f
    # generated values replace the stages; a later copy_in starts from zero
    m.set_random_params(seed=5, alpha=64.0, beta=0.0, scale_kind=0, p0=0.0)
    m.random_matrix()
    f0 = next(f for f in leaves if m.front_info(f)["id"] == 0)
    gen1 = m.fetch_stage(next(f for f in leaves if m.front_info(f)["id"] == 1))
    m.copy_in(E[0], el[0], U.UHM_LHS, K[0])
    om2 = spec.oracle()
    om2.init_numeric(1)
    om2.copy_in(0, [spec.seq[n] for n in el[0]], 1, K[0])
    assert np.array_equal(m.fetch_stage(f0), om2.input(0))
    assert np.array_equal(m.fetch_stage(next(f for f in leaves if m.front_info(f)["id"] == 1)), gen1)
    # and the system (the 4 x 4 golden matrices are unsymmetric: LU) still
    # factors and solves through the device stages
    m.set_rhs()
    assert m.lu_piv_with_free() == 0
    m.solve_lu_piv()
    m.check_lu_piv()
    assert m.backward_error() <= 1e-13
