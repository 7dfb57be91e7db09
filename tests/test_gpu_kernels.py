"""Kernel-level parity of the dominant kernel, the MFMA Schur / trailing
update (uhmk_gemm_update), through the kernel C ABI: every tile size,
operand form and K range the schedule uses, against numpy in FP64, for FP64
products on FP64 fronts, FP32 products on FP32 and FP64 fronts, and BF16
operands with FP32 accumulation on FP32 and FP64 fronts (this also pins the
accumulator lane layout of all three MFMA instructions).  For BF16 the
reference product takes the operands rounded to BF16 (via FP32, nearest
even) and must match to FP32-accumulation accuracy, while differing from the
unrounded product by a BF16-sized amount (the operands really were
rounded)."""
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


# (front storage, MFMA operand type) -> relative tolerance against the FP64
# product of the operands as the MFMA sees them
MODES = {(0, 0): 1e-13, (0, 1): 2e-5, (1, 1): 2e-5, (0, 2): 2e-5, (1, 2): 2e-5}


def bf16_round(x):
    """x rounded once to BF16 (8 significant bits, round to nearest, ties to
    even), as float64 -- the conversion the kernels apply to FP64 and FP32
    operands alike (FP64 values are not rounded to FP32 first).  Values
    within the normal FP32 range."""
    u = np.ascontiguousarray(x, np.float64).view(np.uint64)
    u = (u + np.uint64((1 << 44) - 1) + ((u >> np.uint64(45)) & np.uint64(1))) & ~np.uint64((1 << 45) - 1)
    return u.view(np.float64)


def test_bf16_round_known_values():
    # ties go to the even neighbour: 1 + 2^-8 lies halfway between 1 and
    # 1 + 2^-7, 1 + 3 * 2^-8 halfway between 1 + 2^-7 and 1 + 2^-6; an FP64
    # value just off a tie rounds by its own side, not through FP32 (these
    # six values were also fed through the BF16 MFMA update, which returned
    # exactly these results)
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 1 + 2 ** -8 + 2 ** -40,
                  1 + 2 ** -8 - 2 ** -40, 1 + 3 * 2 ** -8 - 2 ** -40, -2.5])
    want = np.array([1.0, 1.0, 1 + 2 ** -6, -1.0, 1 + 2 ** -7, 1.0, 1 + 2 ** -7, -2.5])
    assert np.array_equal(bf16_round(x), want)
    # any value: 8 significant bits, within half a unit in the last place
    y = np.random.default_rng(0).standard_normal(1000) * 1e30
    r = bf16_round(y)
    assert np.array_equal(r.astype(np.float32).astype(np.float64), r)
    assert not (r.astype(np.float32).view(np.uint32) & 0xFFFF).any()
    assert (np.abs(r - y) <= np.abs(y) * 2.0 ** -8).all()


def _operands(X, mma):
    return bf16_round(X) if mma == 2 else X


def _bf16_really_rounded(G, F, exact, mode):
    """For BF16 operands the result must sit measurably off the unrounded
    product (relative deviation above 1e-4 of the largest entry)."""
    if mode[1] == 2:
        dev = np.abs(G - exact).max() / np.abs(exact).max()
        assert dev > 1e-4, dev


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
    Fo = _operands(F, mode[1])
    ABL = Fo[fs:, :fs]
    B = Fo[fs:, :fs].T if transb else Fo[:fs, fs:]
    want = F.copy()
    want[fs:, fs:] -= ABL @ B
    err = np.abs(G - want).max() / np.abs(want).max()
    assert err < MODES[mode], err
    exact = F.copy()
    exact[fs:, fs:] -= F[fs:, :fs] @ (F[fs:, :fs].T if transb else F[:fs, fs:])
    _bf16_really_rounded(G, F, exact, mode)
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

    def expected(Fo):
        want = F.copy()
        A = Fo[c1:, c0:c1]
        if transb:
            want[c1:, c1:fs] -= A @ Fo[c1:fs, c0:c1].T
        else:
            upd = A @ Fo[c0:c1, c1:]
            mask = np.ones((n - c1, n - c1), bool)
            mask[fs - c1:, fs - c1:] = False  # ABR is deferred to the Schur launch
            want[c1:, c1:][mask] -= upd[mask]
        return want

    want = expected(_operands(F, mode[1]))
    err = np.abs(G - want).max() / np.abs(want).max()
    assert err < MODES[mode], err
    _bf16_really_rounded(G, F, expected(F), mode)
    # entries outside the updated region keep their bits
    assert np.array_equal(G[:, :c1], F[:, :c1]) and np.array_equal(G[:c1, :], F[:c1, :])


@pytest.mark.parametrize("prec", [0, 1])
def test_gemm_bf16_rounding_ties(lib, prec):
    """Operands at and next to BF16 rounding ties through the BF16 update:
    A is a unit selector, so C = -round(B) shows the rounded operand bits."""
    vals = [1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 1 + 2 ** -9, 3.0, -2.5]
    if prec == 0:  # FP64 values just off a tie
        vals += [1 + 2 ** -8 + 2 ** -40, 1 + 2 ** -8 - 2 ** -40, 1 + 3 * 2 ** -8 - 2 ** -40]
    n, fs = 48, 32
    F = np.zeros((n, n))
    for r in range(16):
        F[fs + r, r] = 1.0
    for r, v in enumerate(vals):
        F[r, fs] = v
    print("case bf16 ties prec=%d" % prec, flush=True)
    G = _run(F, fs, [(0, fs, fs, 0)], 0, 0, 1, 0, 64, prec, 2)
    got = -G[fs:fs + len(vals), fs]
    assert np.array_equal(got, bf16_round(np.array(vals))), got
