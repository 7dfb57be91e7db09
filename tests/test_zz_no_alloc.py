"""uhmk_block_dots_noalloc (block dot products with their partial sums in a
fixed device array, for callers that cannot allocate) against
uhmk_block_dots with a caller-allocated workspace, through the kernel C ABI:
bit-identical results on two calls in a row, the second with a smaller
input (the fixed array keeps the first call's partial sums), and a clear
error for sizes above the array."""
import ctypes as C

import numpy as np
import pytest

import uhm_amd as U
from hipmem import DeviceArray, hip

pytestmark = pytest.mark.gpu


def _fns():
    lib = U.load()
    dots = lib.uhmk_block_dots
    dots.restype = C.c_int
    dots.argtypes = [C.c_void_p, C.c_longlong, C.c_int, C.c_void_p, C.c_longlong, C.c_int, C.c_void_p,
                     C.c_void_p, C.c_void_p]
    noalloc = lib.uhmk_block_dots_noalloc
    noalloc.restype = C.c_int
    noalloc.argtypes = [C.c_void_p, C.c_longlong, C.c_int, C.c_void_p, C.c_longlong, C.c_int, C.c_void_p,
                        C.c_void_p]
    scratch = lib.uhmk_block_dots_scratch
    scratch.restype = C.c_longlong
    scratch.argtypes = [C.c_longlong, C.c_int]
    cap = lib.uhmk_block_dots_noalloc_max
    cap.restype = C.c_longlong
    cap.argtypes = []
    return dots, noalloc, scratch, cap


def _case(rng, nrow, nrhs, k):
    V = rng.standard_normal((k, nrhs, nrow))          # k blocks of nrow x nrhs, column major
    w = rng.standard_normal((nrhs, nrow))
    return V, w


def _run_both(dots, noalloc, scratch, V, w):
    k, nrhs, nrow = V.shape
    dV, dw = DeviceArray(V), DeviceArray(w)
    dpart = DeviceArray(np.zeros(max(1, scratch(nrow, nrhs))))
    out_ref = DeviceArray(np.zeros(k * nrhs))
    out_na = DeviceArray(np.zeros(k * nrhs))
    vs = nrow * nrhs
    rc = dots(dV.ptr, vs, k, dw.ptr, nrow, nrhs, dpart.ptr, out_ref.ptr, None)
    assert rc == 0, rc
    rc = noalloc(dV.ptr, vs, k, dw.ptr, nrow, nrhs, out_na.ptr, None)
    assert rc == 0, rc
    ref = out_ref.download().copy()
    got = out_na.download().copy()
    for a in (dV, dw, dpart, out_ref, out_na):
        a.free()
    return ref, got


def test_noalloc_block_dots_match_workspace_version():
    dots, noalloc, scratch, cap = _fns()
    rng = np.random.default_rng(11)
    # first call: 3 row blocks x 32 basis blocks x 5 columns of partial sums;
    # second: fewer rows, columns and blocks, so stale partials of the first
    # call sit beyond what it reads
    for nrow, nrhs, k in ((9000, 5, 32), (5000, 3, 7)):
        assert scratch(nrow, nrhs) <= cap()
        V, w = _case(rng, nrow, nrhs, k)
        print("noalloc block dots: nrow %d nrhs %d k %d" % (nrow, nrhs, k), flush=True)
        ref, got = _run_both(dots, noalloc, scratch, V, w)
        assert np.array_equal(ref.view(np.uint64), got.view(np.uint64))
        exact = np.einsum("icr,cr->ic", V, w).reshape(-1)
        assert np.allclose(got, exact, rtol=1e-12, atol=1e-12 * np.abs(exact).max())


def test_noalloc_block_dots_rejects_oversize():
    _, noalloc, scratch, cap = _fns()
    # sizes whose partial sums exceed the fixed array fail before any launch
    nrow = 4096 * (cap() // 32 + 1)
    assert scratch(nrow, 1) > cap()
    assert noalloc(None, nrow, 1, None, nrow, 1, None, None) == -1013
    assert hip().hipDeviceSynchronize() == 0
