"""Kernel-level parity of the LU panel through the kernel C ABI.

uhmk_lu_panel_ll (left-looking register blocks, rows left in place until the
end) factors one m x 64 panel with partial pivoting restricted to its m rows
(the ATL rows of a front), for every register-block shape it is
instantiated with (m from 40 to 8000, FP64 and FP32 storage).  Checked
against
  * the oracle's unblocked panel (getf2_panel: the reference's FLA_LU_piv on
    the panel, LINAL/src/dense/lu_piv.cxx:79): pivots equal up to the first
    step whose oracle choice was a near tie, L\\U to rounding;
  * the right-looking panel kernel (uhmk_lu_panel), bit for bit on FP64
    fronts up to that kernel's 6144-row limit, including the net row
    permutation the row kernel reads (pperm);
  * the rows and columns outside the panel, which must stay untouched.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import uhm_amd as U
from hipmem import FRONT_DTYPE, DeviceArray
from oracle import oracle

pytestmark = pytest.mark.gpu

PP_STRIDE = 257


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _panel(fn_name, A, fs, c0, nb, prec):
    """Run one panel launch on the n x n front A (column major); returns
    (front, relative pivots of the panel, info, pperm)."""
    n = A.shape[0]
    lib = U.load()
    fn = getattr(lib, fn_name)
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                   C.c_void_p, C.c_void_p, C.c_void_p]
    fr = np.zeros(1, FRONT_DTYPE)
    fr["n"], fr["fs"], fr["parent"], fr["prec"], fr["poff"] = n, fs, -1, prec, 0
    dF = DeviceArray(fr)
    dL = DeviceArray(np.zeros(1, np.int32))
    dA = DeviceArray(np.asfortranarray(A).ravel(order="F").astype(np.float32 if prec else np.float64))
    dP = DeviceArray(np.zeros(max(1, fs), np.int32))
    dI = DeviceArray(np.zeros(1, np.int32))
    dQ = DeviceArray(np.zeros(PP_STRIDE, np.int32))
    rc = fn(dF.ptr, dL.ptr, 1, fs - c0, c0, nb, dA.ptr, prec, dP.ptr, dI.ptr, dQ.ptr, None)
    assert rc == 0, rc
    out = dA.download().reshape(n, n if A.shape[1] == n else A.shape[1], order="F").astype(np.float64)
    kb = min(nb, fs - c0)
    return out, dP.download()[c0:c0 + kb].copy(), int(dI.download()[0]), dQ.download().copy()


def _check_against_oracle(G, piv, A, fs, c0, kb, prec):
    """Pivots of the kernel equal the oracle's up to the first near tie;
    L\\U equal to rounding when every pivot matched."""
    ref, p_ref, marg, _ = oracle.panel_lu(A[c0:fs, c0:c0 + kb], kb)
    tie = 1e-10 if prec == 0 else 1e-4
    agree = 0
    for k in range(kb):
        if piv[k] != p_ref[k]:
            assert marg[k] <= tie, (k, piv[k], p_ref[k], marg[k])
            break
        agree += 1
    if agree == kb:
        got = G[c0:fs, c0:c0 + kb]
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err <= (1e-12 if prec == 0 else 5e-5), err
    return agree


# one case per register-block shape of the kernel (rows per thread R and
# block width W: 64 / 128 / 256 / 512 threads with R = 1, then 512 threads
# with R = 2, 3, 4, 6, 8, 12, 16)
SIZES = [40, 100, 250, 500, 1000, 1500, 2000, 3000, 4000, 6000, 8000]


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("m", SIZES)
def test_panel_vs_oracle(lib, oracle_lib, m, prec):
    rng = np.random.default_rng(1000 + m + prec)
    kb = 64
    A = rng.standard_normal((m, kb))
    if prec:
        A = A.astype(np.float32).astype(np.float64)
    G, piv, info, _ = _panel("uhmk_lu_panel_ll", A, m, 0, kb, prec)
    assert info == 0
    agree = _check_against_oracle(G, piv, A, m, 0, min(kb, m), prec)
    say("  m %d prec %d: %d of %d pivots before the first near tie" % (m, prec, agree, min(kb, m)))


@pytest.mark.parametrize("m", [100, 700, 2100, 4000, 6000])
def test_panel_same_bits_as_right_looking_kernel(lib, m):
    rng = np.random.default_rng(77 + m)
    A = rng.standard_normal((m, 64))
    G1, p1, i1, q1 = _panel("uhmk_lu_panel_ll", A, m, 0, 64, 0)
    G0, p0, i0, q0 = _panel("uhmk_lu_panel", A, m, 0, 64, 0)
    assert np.array_equal(p1, p0)
    assert i1 == i0
    assert np.array_equal(G1.view(np.uint64), G0.view(np.uint64))
    cnt = q0[0]
    assert q1[0] == cnt
    assert np.array_equal(q1[1:1 + cnt], q0[1:1 + cnt])
    assert np.array_equal(q1[129:129 + cnt], q0[129:129 + cnt])


@pytest.mark.parametrize("prec", [0, 1])
def test_panel_inside_front(lib, oracle_lib, prec):
    """A panel at c0 = 64 of a 300 x 300 front with fs = 280 and a short
    last panel at c0 = 256: the ABL rows [fs, n), the rows above c0 and the
    columns outside the panel are left as they were."""
    rng = np.random.default_rng(5 + prec)
    n, fs = 300, 280
    A = rng.standard_normal((n, n))
    if prec:
        A = A.astype(np.float32).astype(np.float64)
    for c0 in (64, 256):
        kb = min(64, fs - c0)
        G, piv, info, _ = _panel("uhmk_lu_panel_ll", A, fs, c0, 64, prec)
        assert info == 0
        _check_against_oracle(G, piv, A, fs, c0, kb, prec)
        outside = np.ones((n, n), bool)
        outside[c0:fs, c0:c0 + kb] = False
        assert np.array_equal(G[outside], A[outside])


def test_panel_zero_column_reported(lib):
    """A zero column gives a zero pivot: info = first zero-pivot row + 1
    (absolute row), and the panel still completes."""
    rng = np.random.default_rng(9)
    n = 600
    A = rng.standard_normal((n, n))
    A[64:, 64 + 5] = 0.0
    G, piv, info, _ = _panel("uhmk_lu_panel_ll", A, n, 64, 64, 0)
    G0, p0, info0, _ = _panel("uhmk_lu_panel", A, n, 64, 64, 0)
    assert info == 64 + 5 + 1 == info0
    assert np.array_equal(piv, p0)
    assert np.array_equal(G.view(np.uint64), G0.view(np.uint64))


def test_panel_over_limit_refused(lib):
    lib_ = U.load()
    fn = lib_.uhmk_lu_panel_ll
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                   C.c_void_p, C.c_void_p, C.c_void_p]
    # the row count is checked before anything is read
    assert fn(None, None, 1, 8193, 0, 64, None, 0, None, None, None, None) == -1001


@pytest.mark.parametrize("kind,dims,p", [(2, (8, 8, 8), 3), (3, (24, 24), 8)])
def test_mesh_same_bits_both_panel_kernels(lib, kind, dims, p):
    """A whole LU factorization with the left-looking panels (tuning key 11
    = 1, the default) equals the one with the right-looking panels bit for
    bit: every front and every pivot (FP64 fronts)."""
    m = U.Mesh()
    m.build_synthetic(kind, dims, p, 12345 if kind == 3 else 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=3, alpha=-1.0 if kind == 3 else 64.0, beta=0.1, scale_kind=1 if kind == 2 else 0,
                        p0=1.0 if kind == 2 else 0.0)
    m.random_matrix()
    m.set_rhs()
    assert m.get_tuning(11) == 1
    m.set_tuning(11, 0)
    assert m.lu_piv_with_free() == 0
    F0 = [m.fetch_front(f) for f in range(m.n_fronts())]
    P0 = [m.fetch_pivots(f) for f in range(m.n_fronts())]
    m.set_tuning(11, 1)
    assert m.lu_piv_with_free() == 0
    tall = 0
    for f in range(m.n_fronts()):
        tall += m.front_info(f)["fs"] > 64
        assert np.array_equal(P0[f], m.fetch_pivots(f)), f
        assert np.array_equal(F0[f].view(np.uint64), m.fetch_front(f).view(np.uint64)), f
    say("  fronts with more than one panel:", tall)
    assert tall > 0
    m.solve_lu_piv()
    m.check_lu_piv()
    assert m.backward_error() <= 1e-13


@pytest.mark.parametrize("prec", [U.PREC_FP64, U.PREC_MIXED])
def test_lu_lookahead_same_bits(lib, prec):
    """The LU lookahead (tuning key 12: the next panel on a second stream
    beside the rest of the update) computes every tile and row item as the
    in-order schedule does, so the factors and pivots are the same bits,
    also when the schedule is captured in a graph and replayed."""
    m = U.Mesh()
    m.build_synthetic(2, (9, 8, 8), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=21, alpha=64.0, beta=0.1, scale_kind=1, p0=1.0)
    m.random_matrix()
    m.set_rhs()
    m.set_precision(prec, 2)
    m.set_tuning(12, 0)
    assert m.lu_piv_with_free() == 0
    F0 = [m.fetch_front(f) for f in range(m.n_fronts())]
    P0 = [m.fetch_pivots(f) for f in range(m.n_fronts())]
    m.set_tuning(12, 1)
    for graph in (False, True, True):
        m.set_graph(graph)
        assert m.lu_piv_with_free() == 0
        for f in range(m.n_fronts()):
            assert np.array_equal(P0[f], m.fetch_pivots(f)), (graph, f)
            assert np.array_equal(F0[f].view(np.uint64), m.fetch_front(f).view(np.uint64)), (graph, f)
    m.set_graph(False)
    m.solve_lu_piv()
    m.check_lu_piv()
    if prec == U.PREC_FP64:
        assert m.backward_error() <= 1e-13


def test_lu_row_chunks_same_bits(lib):
    """The row kernel (swaps, U strip, ABL rows) in chunks of 16, 32 or 64
    columns (tuning key 13) gives every column and row the same operations:
    the same factors bit for bit."""
    m = U.Mesh()
    m.build_synthetic(2, (8, 9, 8), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=5, alpha=64.0, beta=0.1, scale_kind=1, p0=1.0)
    m.random_matrix()
    m.set_rhs()
    ref = None
    for w in (64, 32, 16):
        m.set_tuning(13, w)
        assert m.lu_piv_with_free() == 0
        F = [m.fetch_front(f) for f in range(m.n_fronts())]
        P = [m.fetch_pivots(f) for f in range(m.n_fronts())]
        if ref is None:
            ref = (F, P)
            continue
        for f in range(m.n_fronts()):
            assert np.array_equal(ref[1][f], P[f]), (w, f)
            assert np.array_equal(ref[0][f].view(np.uint64), F[f].view(np.uint64)), (w, f)
    m.solve_lu_piv()
    m.check_lu_piv()
    assert m.backward_error() <= 1e-13

