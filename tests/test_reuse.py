"""Factorization reuse (Mesh_::backup / check_reuse,
UHM/src/mesh/mesh.cxx:113-225, and the skip in op_decompose,
UHM/src/operation/element.cxx:204-206).

CPU: the reuse flags.  An unchanged mesh keeps every element; a new dof
count on the interior (factor) node of one leaf clears that leaf and all
its ancestors -- whose own node lists did not change, so only the rule
"cannot reuse upper hierarchy" clears them; a shared (schur) node clears
every element that holds it and their ancestors; elements the backup does
not know are not reusable.

GPU: after such a change, the factorization that keeps the reusable fronts
gives every front and the solution bit for bit as a fresh mesh with the same
change factored from scratch, with fewer flops."""
import numpy as np
import pytest

import uhm_amd as U


def _mesh(dims=(3, 3, 3), p=3):
    m = U.Mesh()
    m.build_synthetic(2, dims, p, 0)
    return m


def _ancestors(m, f):
    out = []
    while True:
        f = m.front_info(f)["parent"]
        if f < 0:
            return out
        out.append(f)


def _flags(m, bk):
    """Front indices whose element is flagged reusable, through the public
    count and a per-front probe (changing nothing)."""
    m.check_reuse(bk)
    return m.n_reusable()


def test_reuse_flags_follow_the_hierarchy(lib):
    m = _mesh()
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    nf = m.n_fronts()
    bk = U.Mesh()
    m.backup(bk)
    m.check_reuse(bk)
    assert m.n_reusable() == nf
    # a leaf's interior node: the leaf and its ancestors only
    leaf = next(f for f in range(nf) if m.front_info(f)["leaf"] and m.front_info(f)["fs"] > 0)
    nid = int(m.front_nodes(leaf, 0)[0][0])
    m.unlock()
    m.set_node(nid, 4, 2)
    m.lock()
    m.check_reuse(bk)
    assert m.n_reusable() == nf - 1 - len(_ancestors(m, leaf))
    # back to the old values (synthetic hexahedra: a node per grid point,
    # one dof, p = 3): reusable again (same node, same code)
    m.unlock()
    m.set_node(nid, 1, 3)
    m.lock()
    m.check_reuse(bk)
    assert m.n_reusable() == nf
    # a shared node: every element holding it (as factor or schur node) and
    # every ancestor of those
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    sid = int(m.front_nodes(leaf, 1)[0][0])
    holders = {f for f in range(nf) for w in (0, 1) if sid in set(m.front_nodes(f, w)[0].tolist())}
    cleared = set(holders)
    for f in holders:
        cleared.update(_ancestors(m, f))
    m.unlock()
    m.set_node(sid, 1, 1)
    m.lock()
    m.check_reuse(bk)
    assert len(holders) >= 2
    assert m.n_reusable() == nf - len(cleared)
    # a backup of another mesh knows none of these nodes
    other = _mesh((2, 2, 2))
    other.lock()
    bk2 = U.Mesh()
    other.backup(bk2)
    m.check_reuse(bk2)
    assert m.n_reusable() == 0
    with pytest.raises(U.UHMError):
        m.set_node(nid, 4, 2)  # locked


def _factored(m, seed, prec):
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    if prec:
        m.set_precision(prec, 1)  # a property of the fronts: set again for every new pattern
    m.create_matrix_buffer()
    m.set_random_params(seed=seed, alpha=64.0, beta=0.0, scale_kind=1, p0=0.5)
    m.random_matrix()
    m.set_rhs()
    m.set_profile(True)
    assert m.chol_with_free() == 0
    flops = sum(v[1] for v in m.kernel_stats().values())
    m.solve_chol()
    m.check_chol()
    return flops


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [U.PREC_FP64, U.PREC_MIXED])
def test_refactorization_keeps_reusable_fronts(lib, prec):
    print("case reuse, precision mode", prec, flush=True)
    m = _mesh((4, 4, 4))
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    leaf = next(f for f in range(m.n_fronts()) if m.front_info(f)["leaf"] and m.front_info(f)["fs"] > 0)
    nid = int(m.front_nodes(leaf, 0)[0][0])
    _factored(m, 7, prec)
    bk = U.Mesh()
    m.backup(bk)
    m.unlock()
    m.set_node(nid, 4, 2)
    m.lock()
    m.check_reuse(bk)
    keep = m.n_reusable()
    flops = _factored(m, 7, prec)
    assert m.n_reused() == keep > 0
    eta = m.backward_error()
    # the same change on a fresh mesh, factored from scratch
    f2 = _mesh((4, 4, 4))
    f2.set_node(nid, 4, 2)
    f2.lock()
    full = _factored(f2, 7, prec)
    print("  kept %d of %d fronts, flops %.3g vs %.3g, eta %.2e" % (keep, m.n_fronts(), flops, full, eta), flush=True)
    assert m.n_fronts() == f2.n_fronts()
    for f in range(m.n_fronts()):
        fs = m.front_info(f)["fs"]
        A, B = m.fetch_front(f), f2.fetch_front(f)
        # the factor as stored: the lower triangle and the 64 x 64 diagonal
        # blocks of ATL (their strict upper parts hold the inverses); the
        # rest of the upper triangle is never written
        keepm = np.tril(np.ones(A.shape, bool))
        for b0 in range(0, fs, 64):
            b1 = min(b0 + 64, fs)
            keepm[b0:b1, b0:b1] = True
        assert np.array_equal(A.view(np.uint64)[keepm], B.view(np.uint64)[keepm]), f
    assert np.array_equal(m.fetch_x(), f2.fetch_x())
    assert flops < full
    assert 0.0 < eta <= (1e-13 if not prec else 1e-5)
