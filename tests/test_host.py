"""Host-side logic of the product against the oracle (CPU only): the lock
analysis, numbering, mapper runs and front sizes are integer parity targets
and must match bit for bit.  Also checks that the C ABI library loads and
exports every symbol the headers declare, and that the numeric path refuses
to run without a GPU (no CPU fallback)."""
import os
import re

import numpy as np
import pytest

import uhm_amd as U
from meshes import MeshSpec, ex1_spec, qp_spec, spec_from_connectivity

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(uhmk?_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(lib):
    names = _declared("uhm_amd.h") + _declared("uhmk.h")
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes table covers the whole uhm_amd.h surface
    assert set(_declared("uhm_amd.h")) <= set(U.exported_symbols())


def _compare(m, spec, om):
    nf = m.n_fronts()
    assert nf == spec.n_elem
    seen = set()
    for f in range(nf):
        info = m.front_info(f)
        e = info["id"]
        seen.add(e)
        el = om.element(e)
        assert (info["fs"], info["ss"], info["generation"]) == (el["fs"], el["ss"], el["generation"]), e
        ids, offs = m.front_nodes(f, 0)
        assert ids.tolist() == [spec.nodes[k][0] for k in el["factor"]]
        assert offs.tolist() == el["factor_off"].tolist()
        ids, offs = m.front_nodes(f, 1)
        assert ids.tolist() == [spec.nodes[k][0] for k in el["schur"]]
        assert offs.tolist() == el["schur_off"].tolist()
        assert m.mapper(f).tolist() == om.mapper(e).tolist()
    assert seen == set(range(spec.n_elem))
    for k, (nid, _) in enumerate(spec.nodes):
        assert m.node_offset(nid) == om.node_offset(k)
    # schedule: levels leaves first, element id order inside a level
    lv = [m.front_info(f)["level"] for f in range(nf)]
    assert lv == sorted(lv)
    oracle_order = [int(e) for L in om.levels() for e in L]
    assert [m.front_info(f)["id"] for f in range(nf)] == oracle_order


def test_lock_ex1_matches_hand_derivation_and_oracle(lib, oracle_lib):
    spec, g = ex1_spec()
    m, E = spec.build_product(U)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    assert m.get_n_dof() == g["ndof"]
    for f in range(3):
        info = m.front_info(f)
        e = info["id"]
        assert info["fs"] == g["fs"][e] and info["ss"] == g["ss"][e]
        ids, offs = m.front_nodes(f, 0)
        assert [[int(a), int(b)] for a, b in zip(ids, offs)] == g["factor"][e]
        assert m.mapper(f).tolist() == g["mapper"][e]
    for nid, o in g["offset"].items():
        assert m.node_offset(int(nid)) == o
    _compare(m, spec, spec.oracle())


@pytest.mark.parametrize("dims,p", [((4, 4), 2), ((16, 16), 2), ((3, 2, 2), 2), ((4, 4, 4), 3), ((7, 5), 3)])
def test_lock_api_mesh_matches_oracle(lib, oracle_lib, dims, p):
    spec = qp_spec(dims, p)
    m, _ = spec.build_product(U)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    _compare(m, spec, spec.oracle())


@pytest.mark.parametrize("kind,dims,p,seed", [(1, (16, 16), 2, 0), (2, (6, 5, 4), 3, 0), (3, (24, 20), 8, 12345),
                                              (3, (9, 13), 4, 5)])
def test_lock_synthetic_mesh_matches_oracle(lib, oracle_lib, kind, dims, p, seed):
    m = U.Mesh()
    m.build_synthetic(kind, dims, p, seed)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    spec = spec_from_connectivity(m.export_connectivity())
    if kind != 3:
        ref = qp_spec(dims, p)
        assert spec.elems == ref.elems and spec.parent == ref.parent and spec.children == ref.children
    _compare(m, spec, spec.oracle())


@pytest.mark.parametrize("kind,dims,p,ndof,root", [(1, (16, 16), 2, 1089, 33), (2, (16, 16, 16), 3, 117649, 2401),
                                                   (2, (16, 16, 16), 4, 274625, 4225),
                                                   (2, (24, 24, 24), 3, 389017, 5329)])
def test_config_sizes(lib, kind, dims, p, ndof, root):
    """dof counts and root separators of the benchmark configurations."""
    m = U.Mesh()
    m.build_synthetic(kind, dims, p, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    assert m.get_n_dof() == ndof
    roots = [f for f in range(m.n_fronts()) if m.front_info(f)["parent"] < 0]
    assert len(roots) == 1 and m.front_info(roots[0])["fs"] == root and m.front_info(roots[0])["ss"] == 0


def test_flops_match_reference_formula(lib):
    m = U.Mesh()
    m.build_synthetic(2, (16, 16, 16), 3, 0)
    m.lock()
    fd, _, _, _ = m.estimate_cost(U.UHM_LU_PIV)
    assert abs(fd / 1e9 - 177.777) < 1e-3
    m = U.Mesh()
    m.build_synthetic(2, (16, 16, 16), 4, 0)
    m.lock()
    fd, _, _, _ = m.estimate_cost(U.UHM_CHOL)
    assert abs(fd / 1e9 - 494.846) < 1e-3


def test_copy_in_checks_dimensions(lib):
    spec = qp_spec((2, 2), 1)
    m, E = spec.build_product(U)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    with pytest.raises(U.UHMError, match="m does not match"):
        m.copy_in(E[0], spec.elems[0], U.UHM_LHS, np.eye(3))
    with pytest.raises(U.UHMError, match="cannot find node"):
        m.copy_in(E[0], [999, 0, 1, 2], U.UHM_LHS, np.eye(4))


def test_lock_requires_tree_and_rejects_late_nodes(lib):
    spec = qp_spec((2, 1), 1)
    m, E = spec.build_product(U)
    m.lock()
    with pytest.raises(U.UHMError, match="locked"):
        m.add_node(10_000, 1)


def test_numeric_path_fails_loudly_without_gpu(lib):
    import ctypes
    try:
        n = ctypes.c_int(0)
        ok = ctypes.CDLL("libamdhip64.so").hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        ok = False
    if ok:
        pytest.skip("a GPU is present")
    spec = qp_spec((2, 2), 1)
    m, E = spec.build_product(U)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    with pytest.raises(U.UHMError, match="HIP device|GPU"):
        m.lu_piv_with_free()


def test_import_uhm_fixture_then_build_tree_matches_oracle_lock(lib, oracle_lib):
    """uhm_mesh_import_ (mesh.cxx:511-579) on a committed ".uhm" file: the
    mesh is emptied first, the nodes no element holds are dropped
    (remove_orphan_nodes, mesh.cxx:246-251, 573) and the others keep their
    creation order; the leaf elements are the file's.  uhm_build_tree_ then
    gives a tree whose lock (fronts, numbering, mapper runs, schedule)
    equals the oracle's lock of the same connectivity."""
    path = os.path.join(ROOT, "tests", "golden", "quad_q2_6x5.uhm")
    spec = qp_spec((6, 5), 2)
    m = U.Mesh()
    m.build_synthetic(1, (2, 2), 2, 0)  # content that the import replaces
    m.import_file(path)
    c = m.export_connectivity()
    assert c["node_id"].tolist() == [nid for nid, _ in spec.nodes]
    assert len(c["parent"]) == len(spec.elems) and (c["parent"] == -1).all()
    ids = c["node_id"]
    for e, nl in enumerate(spec.elems):
        got = [int(ids[k]) for k in c["enode"][c["eptr"][e]:c["eptr"][e + 1]]]
        assert sorted(got) == sorted(nl), e
    m.build_tree()
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    c = m.export_connectivity()
    roots = [e for e, p in enumerate(c["parent"]) if p < 0]
    assert len(roots) == 1 and len(c["parent"]) == 2 * len(spec.elems) - 1
    assert m.get_n_dof() == len(spec.nodes)
    tree = spec_from_connectivity(c)
    _compare(m, tree, tree.oracle())


def _chain_mesh(n_elem, first_id=0, m=None):
    """1D chain: element k holds end nodes n_k, n_{k+1} and interior node
    i_k (one dof each)."""
    m = m or U.Mesh()
    ends = [m.add_node(first_id + k, 1) for k in range(n_elem + 1)]
    inner = [m.add_node(first_id + 1000 + k, 1) for k in range(n_elem)]
    for k in range(n_elem):
        e = m.add_element()
        for nd in (ends[k], inner[k], ends[k + 1]):
            e.add_node(nd)
    return m


def test_build_tree_coarsening_hand_derived_chain(lib):
    """build_tree_var_1 (build_tree_var1.cxx:53-181) on a chain of four
    elements, derived by hand: the shared end nodes n1..n3 are the schur
    nodes; branch weights (shared dofs, minus other schur dofs) are
    (e0,e1) = (1,-1), (e1,e2) = (1,-2), (e2,e3) = (1,-1), visited from the
    largest, equal weights last-inserted first: (e2,e3) -> element 4,
    (e0,e1) -> element 5, (e1,e2) skipped; then (4,5) -> root 6."""
    m = _chain_mesh(4)
    m.build_tree()
    c = m.export_connectivity()
    assert c["parent"].tolist() == [5, 5, 4, 4, 6, 6, -1]
    kids = [c["cidx"][c["cptr"][e]:c["cptr"][e + 1]].tolist() for e in range(7)]
    assert kids == [[], [], [], [], [2, 3], [0, 1], [4, 5]]
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    assert m.get_n_dof() == 9


def test_build_tree_coarsening_mesh_matches_oracle_lock(lib, oracle_lib):
    """On a 2D Q2 mesh of leaves only: one binary tree over all elements,
    generations decreasing to the root, and its lock equal to the oracle's
    lock of the same connectivity."""
    spec = qp_spec((7, 6), 2)
    leaves = MeshSpec(spec.nodes, spec.elems, [-1] * len(spec.elems), [[] for _ in spec.elems])
    m, _ = leaves.build_product(U)
    m.build_tree()
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    c = m.export_connectivity()
    ne = len(spec.elems)
    assert len(c["parent"]) == 2 * ne - 1
    assert sum(1 for p in c["parent"] if p < 0) == 1
    for e in range(ne, 2 * ne - 1):
        assert c["cptr"][e + 1] - c["cptr"][e] == 2
    assert m.get_n_dof() == len(spec.nodes)
    tree = spec_from_connectivity(c)
    _compare(m, tree, tree.oracle())


def test_build_tree_disconnected_mesh_single_root(lib):
    """Two chains with no node in common never pair by coarsening (the
    reference's loop would not end); their roots are joined into one tree."""
    m = _chain_mesh(3)
    _chain_mesh(3, first_id=500, m=m)
    m.build_tree()
    c = m.export_connectivity()
    assert len(c["parent"]) == 2 * 6 - 1
    assert sum(1 for p in c["parent"] if p < 0) == 1
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    assert m.get_n_dof() == 2 * (4 + 3)
