"""Kernel-level parity of the dominant kernel, the MFMA Schur / trailing
update (uhmk_gemm_update), through the kernel C ABI: every tile size,
operand form and K range the schedule uses, against numpy in FP64, for FP64
products on FP64 fronts and FP32 products on FP32 and FP64 fronts (this
also pins the accumulator lane layout of both MFMA instructions)."""
import ctypes as C
import sys

import numpy as np
import pytest

import uhm_amd as U
from hipmem import FRONT_DTYPE, DeviceArray

pytestmark = pytest.mark.gpu


def _tiles(a, b, cut, T):
    out = list(range(a, min(b, cut), T))
    out += list(range(max(a, cut), b, T))
    return out


# (front storage, MFMA operand type) -> relative tolerance against FP64
MODES = {(0, 0): 1e-13, (0, 1): 2e-5, (1, 1): 2e-5}


def _run(F, fs, items, c0, nb, schur, transb, tile, prec=0, mma=0):
    n = F.shape[0]
    lib = U.load()
    fn = lib.uhmk_gemm_update
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                   C.c_int, C.c_int, C.c_void_p]
    fr = np.zeros(1, FRONT_DTYPE)
    fr["n"], fr["fs"], fr["parent"], fr["prec"] = n, fs, -1, prec
    dF = DeviceArray(fr)
    dI = DeviceArray(np.array(items, np.int32).reshape(-1, 4))
    dA = DeviceArray(np.asfortranarray(F).ravel(order="F").astype(np.float32 if prec else np.float64))
    rc = fn(dF.ptr, dI.ptr, len(items), c0, nb, schur, transb, tile, dA.ptr, prec, mma, None)
    assert rc == 0
    return dA.download().reshape(n, n, order="F").astype(np.float64)


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("tile", [64, 128])
@pytest.mark.parametrize("transb", [0, 1])
def test_gemm_schur_tiles(lib, tile, transb, mode):
    rng = np.random.default_rng(5 + tile + transb)
    n, fs = 421, 157  # ragged: neither n nor fs is a tile multiple
    F = rng.standard_normal((n, n))
    if mode[0]:
        F = F.astype(np.float32).astype(np.float64)
    items = [(0, i0, j0, 0) for j0 in range(fs, n, tile) for i0 in range(fs, n, tile)]
    print("case schur tile=%d transb=%d mode=%s" % (tile, transb, mode), flush=True)
    G = _run(F, fs, items, 0, 0, 1, transb, tile, *mode)
    ABL = F[fs:, :fs]
    B = F[fs:, :fs].T if transb else F[:fs, fs:]
    want = F.copy()
    want[fs:, fs:] -= ABL @ B
    err = np.abs(G - want).max() / np.abs(want).max()
    assert err < MODES[mode], err
    # nothing outside ABR moves
    assert np.array_equal(G[:, :fs], F[:, :fs]) and np.array_equal(G[:fs, :], F[:fs, :])


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("tile", [64, 128])
@pytest.mark.parametrize("transb", [0, 1])
def test_gemm_panel_update(lib, tile, transb, mode):
    rng = np.random.default_rng(11 + tile + transb)
    n, fs, c0, nb = 399, 230, 64, 64
    c1 = min(c0 + nb, fs)
    F = rng.standard_normal((n, n))
    if mode[0]:
        F = F.astype(np.float32).astype(np.float64)
    R = _tiles(c1, n, fs, tile)
    Cc = _tiles(c1, fs if transb else n, fs, tile)
    items = []
    for j0 in Cc:
        for i0 in R:
            if not transb and i0 >= fs and j0 >= fs:
                continue
            items.append((0, i0, j0, 0))
    print("case panel tile=%d transb=%d mode=%s" % (tile, transb, mode), flush=True)
    G = _run(F, fs, items, c0, nb, 0, transb, tile, *mode)
    want = F.copy()
    A = F[c1:, c0:c1]
    if transb:
        want[c1:, c1:fs] -= A @ F[c1:fs, c0:c1].T
    else:
        upd = A @ F[c0:c1, c1:]
        mask = np.ones((n - c1, n - c1), bool)
        mask[fs - c1:, fs - c1:] = False  # ABR is deferred to the Schur launch
        want[c1:, c1:][mask] -= upd[mask]
    err = np.abs(G - want).max() / np.abs(want).max()
    assert err < MODES[mode], err
    # entries outside the updated region keep their bits
    assert np.array_equal(G[:, :c1], F[:, :c1]) and np.array_equal(G[:c1, :], F[:c1, :])
