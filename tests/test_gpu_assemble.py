"""Extend-add parity, bit for bit: uhmk_assemble (gather form: every parent
entry is ((0 + own stage) + C1[inv1]) + C2[inv2] ..., children in merge
order) against the oracle's merge_A (UHM/src/matrix/uhm/fla/matrix.cxx:284-303
through Helper_::_merge_A, helper.hxx:191-222: a zeroed parent, or its own
copy-in, plus each child's ABR scattered through the mapper runs, children
in insertion order), through the kernel C ABI.

The children are the fronts the GPU factored on a real mesh (so their ABR
blocks are genuine Schur complements), copied into a fresh device arena next
to the parent; the parent's maps come from the mesh's own mapper runs.  With
FP32 children (mixed precision) the sums still run in FP64 and are rounded
once to the parent's storage type, as the oracle's FP64 sum rounded once."""
import ctypes as C

import numpy as np
import pytest

import uhm_amd as U
from hipmem import FRONT_DTYPE, DeviceArray
from oracle.oracle import OracleMesh

pytestmark = pytest.mark.gpu


def _assemble_fn():
    fn = U.load().uhmk_assemble
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                   C.c_int, C.c_void_p, C.c_void_p]
    return fn


def _gpu_assemble(m, P, kids, fronts_of, stage, prec_of):
    """Assemble front P from its own stage (leaf) or its children's factored
    fronts in a fresh arena; returns the parent as float64."""
    fiP = fronts_of[P]
    nP, fsP = fiP["fs"] + fiP["ss"], fiP["fs"]
    members = [P] + kids
    fr = np.zeros(len(members), FRONT_DTYPE)
    a64, a32 = [], []
    off64 = off32 = 0
    for q, f in enumerate(members):
        fi = fronts_of[f]
        n = fi["fs"] + fi["ss"]
        vals = np.zeros((n, n)) if q == 0 else m.fetch_front(f)
        p = prec_of[f]
        fr[q]["n"], fr[q]["fs"], fr[q]["prec"], fr[q]["parent"], fr[q]["soff"] = n, fi["fs"], p, -1, -1
        if p:
            fr[q]["off"] = off32
            a32.append(vals.ravel(order="F").astype(np.float32))
            off32 += n * n
        else:
            fr[q]["off"] = off64
            a64.append(vals.ravel(order="F"))
            off64 += n * n
    srcs, inv = [], []
    if kids:
        for q, c in enumerate(kids):
            iv = np.full(nP, -1, np.int32)
            for offs_c, offs_p, fs_p, nd in m.mapper(c):
                base = (fsP if fs_p else 0) + offs_p
                iv[base:base + nd] = np.arange(offs_c, offs_c + nd)
            assert len(set(iv[iv >= 0].tolist())) == fronts_of[c]["ss"]
            srcs.append((q + 1, q * nP))
            inv.append(iv)
    else:
        fr[0]["soff"] = 0
        srcs.append((-1, -1))
    src_beg = np.array([0] + [len(srcs)] * len(members), np.int32)
    items = np.array([(0, i0, j0, 0) for j0 in range(0, nP, 64) for i0 in range(0, nP, 64)], np.int32)
    d = dict(F=DeviceArray(fr), I=DeviceArray(items), sb=DeviceArray(src_beg),
             S=DeviceArray(np.array(srcs, np.int32).reshape(-1, 2)),
             inv=DeviceArray(np.concatenate(inv) if inv else np.zeros(1, np.int32)),
             a64=DeviceArray(np.concatenate(a64) if a64 else np.zeros(1)),
             a32=DeviceArray(np.concatenate(a32) if a32 else np.zeros(1, np.float32)),
             st=DeviceArray(stage.ravel(order="F") if stage is not None else np.zeros(1)))
    rc = _assemble_fn()(d["F"].ptr, d["I"].ptr, len(items), d["sb"].ptr, d["S"].ptr, d["inv"].ptr, d["a64"].ptr,
                        d["a32"].ptr, prec_of[P], d["st"].ptr, None)
    assert rc == 0, rc
    out = (d["a32"] if prec_of[P] else d["a64"]).download()[:nP * nP].astype(np.float64)
    for a in d.values():
        a.free()
    return out.reshape(nP, nP, order="F")


@pytest.mark.parametrize("method,prec", [(U.UHM_LU_PIV, U.PREC_FP64), (U.UHM_CHOL, U.PREC_MIXED)])
def test_assemble_bitwise_equal_to_merge_A(lib, oracle_lib, method, prec):
    print("case assemble method %d prec %d" % (method, prec), flush=True)
    m = U.Mesh()
    m.build_synthetic(2, (4, 4, 3), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=31, alpha=64.0, beta=0.1 if method == U.UHM_LU_PIV else 0.0, scale_kind=1, p0=1.0)
    m.random_matrix()
    m.set_precision(prec, 1)
    assert (m.chol_with_free() if method == U.UHM_CHOL else m.lu_piv_with_free()) == 0
    nf = m.n_fronts()
    fronts_of = [m.front_info(f) for f in range(nf)]
    top = max(fi["level"] for fi in fronts_of)
    # storage per level as the mixed path lays it out: FP64 for the top level
    prec_of = [0 if (prec == U.PREC_FP64 or fi["level"] > top - 1) else 1 for fi in fronts_of]
    front_of_elem = {fi["id"]: f for f, fi in enumerate(fronts_of)}
    c = m.export_connectivity()
    om = OracleMesh.from_connectivity(c)
    om.init_numeric(1)
    stages = {}
    for f, fi in enumerate(fronts_of):
        if fi["leaf"]:
            stages[f] = m.fetch_stage(f)
            om.set_front(fi["id"], stages[f])
    factored = {}
    for f, fi in enumerate(fronts_of):
        vals = m.fetch_front(f)
        factored[f] = vals.astype(np.float32).astype(np.float64) if prec_of[f] else vals
        om.set_factored(fi["id"], factored[f])
    checked = 0
    parents = [f for f in range(nf) if not fronts_of[f]["leaf"]] + [f for f in range(nf) if fronts_of[f]["leaf"]][:4]
    for P in parents:
        e = fronts_of[P]["id"]
        kids = [front_of_elem[int(k)] for k in c["cidx"][c["cptr"][e]:c["cptr"][e + 1]]]
        got = _gpu_assemble(m, P, kids, fronts_of, stages.get(P), prec_of)
        want = om.assemble(e)
        om.set_factored(e, factored[P])  # assemble() overwrote P's values; P is a child further up
        if prec_of[P]:
            want = want.astype(np.float32).astype(np.float64)
        assert np.array_equal(got, want), (P, np.abs(got - want).max())
        checked += 1
    assert checked > 10
