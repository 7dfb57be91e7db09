"""Low-precision Schur complement products through the kernel C ABI
(uhmk_schur_lp_pack + uhmk_schur_lp): the packed BF16 / FP16 operands and
the FP32-accumulated update of ABR, against numpy.

The reference product takes the operands rounded as the kernels round them
(BF16: nearest even via the bit pattern; FP16: numpy's float16 cast of the
operand times its power-of-two scale, the largest entry of each front's
operand in [2^14, 2^15)) and must match to FP32-accumulation accuracy,
while the rounding must be visible against the unrounded product.  Cases:
LU (ABL ATR) and Cholesky (ABL ABL^T, lower tiles), FP32 and FP64 fronts,
tiles of 128 and 64, ragged fs and ss, and operands of magnitude 1e6 and
1e-6 for FP16 (outside FP16's range without the scale)."""
import ctypes as C

import numpy as np
import pytest

import uhm_amd as U
from hipmem import FRONT_DTYPE, DeviceArray

pytestmark = pytest.mark.gpu


def bf16_round(x):
    u = np.ascontiguousarray(x, np.float64).view(np.uint64)
    u = (u + np.uint64((1 << 44) - 1) + ((u >> np.uint64(45)) & np.uint64(1))) & ~np.uint64((1 << 45) - 1)
    return u.view(np.float64)


def fp16_scaled(X):
    """X rounded to FP16 after the power-of-two scale, and that scale."""
    amax = float(np.abs(X).astype(np.float32).max())
    if amax == 0.0:
        return X.copy(), 1.0
    _, e = np.frexp(np.float32(amax))
    s = 2.0 ** (15 - int(e))
    return (X * s).astype(np.float16).astype(np.float64), s


def _fns():
    lib = U.load()
    size = lib.uhmk_schur_lp_size
    size.restype = C.c_longlong
    size.argtypes = [C.c_int, C.c_int, C.c_int]
    pack = lib.uhmk_schur_lp_pack
    pack.restype = C.c_int
    pack.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                     C.c_void_p, C.c_void_p]
    prod = lib.uhmk_schur_lp
    prod.restype = C.c_int
    prod.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                     C.c_void_p, C.c_void_p, C.c_void_p]
    return size, pack, prod


def _run(F, fs, lu, prec, fp16, tile):
    size, pack, prod = _fns()
    n = F.shape[0]
    ss = n - fs
    fr = np.zeros(1, FRONT_DTYPE)
    fr["n"], fr["fs"], fr["parent"], fr["prec"] = n, fs, -1, prec
    pitems = [(0, a, b, op) for op in range(2 if lu else 1) for a in range(((ss + 127) // 128) * 2)
              for b in range((fs + 63) // 64)]
    sitems = [(0, i0, j0, 0) for j0 in range(fs, n, tile) for i0 in range(fs, n, tile) if lu or i0 + tile > j0]
    dF = DeviceArray(fr)
    dP = DeviceArray(np.array(pitems, np.int32).reshape(-1, 4))
    dS = DeviceArray(np.array(sitems, np.int32).reshape(-1, 4))
    dA = DeviceArray(np.asfortranarray(F).ravel(order="F").astype(np.float32 if prec else np.float64))
    doff = DeviceArray(np.zeros(1, np.int64))
    damax = DeviceArray(np.zeros(2, np.uint32))
    dlp = DeviceArray(np.zeros(max(1, size(n, fs, 1 if lu else 0)), np.uint16))
    assert pack(dF.ptr, dP.ptr, len(pitems), fp16, dA.ptr, prec, doff.ptr, damax.ptr, dlp.ptr, None) == 0
    assert prod(dF.ptr, dS.ptr, len(sitems), 1 if lu else 0, tile, fp16, dA.ptr, prec, doff.ptr, damax.ptr, dlp.ptr,
                None) == 0
    G = dA.download().reshape(n, n, order="F").astype(np.float64)
    lp = dlp.download().copy()
    return G, lp


CASES = [
    # name, n, fs, lu, front storage, fp16, tile, magnitude
    ("lu-bf16-f32-128", 421, 157, True, 1, 0, 128, 1.0),
    ("lu-bf16-f64-64", 333, 75, True, 0, 0, 64, 1.0),
    ("chol-bf16-f32-128", 450, 130, False, 1, 0, 128, 1.0),
    ("chol-bf16-f64-128", 300, 64, False, 0, 0, 128, 1.0),
    ("lu-fp16-f32-128", 421, 157, True, 1, 1, 128, 1e6),
    ("chol-fp16-f64-64", 350, 100, False, 0, 1, 64, 1e-6),
    ("chol-fp16-f32-128", 450, 200, False, 1, 1, 128, 1.0),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_lowprec_schur_matches_rounded_product(lib, case):
    name, n, fs, lu, prec, fp16, tile, mag = case
    print("case", name, flush=True)
    rng = np.random.default_rng(n + fs)
    F = rng.standard_normal((n, n)) * mag
    # ABR starts at zero, so every comparison below is relative to the product
    F[fs:, fs:] = 0.0
    if prec:
        F = F.astype(np.float32).astype(np.float64)
    G, lp = _run(F, fs, lu, prec, fp16, tile)
    ABL = F[fs:, :fs]
    ATR = F[:fs, fs:]
    if fp16:
        A, sa = fp16_scaled(ABL)
        B, sb = fp16_scaled(ATR.T if lu else ABL)
        P = (A @ B.T) / (sa * sb)
    else:
        A, B = bf16_round(ABL), bf16_round(ATR.T if lu else ABL)
        P = A @ B.T
    want = F.copy()
    want[fs:, fs:] -= P
    region = np.ones((n - fs, n - fs), bool) if lu else np.tril(np.ones((n - fs, n - fs), bool))
    scale = np.abs(want[fs:, fs:][region]).max()
    err = np.abs(G[fs:, fs:] - want[fs:, fs:])[region].max() / scale
    exact = F[fs:, fs:] - ABL @ (ATR if lu else ABL.T)
    dev = np.abs(G[fs:, fs:] - exact)[region].max() / scale
    print("  err vs rounded product %.2e, vs exact product %.2e" % (err, dev), flush=True)
    assert err < 2e-5
    assert dev > (1e-6 if fp16 else 1e-5)  # the operands really were rounded
    # nothing outside ABR moves
    assert np.array_equal(G[:, :fs], F[:, :fs]) and np.array_equal(G[:fs, :], F[:fs, :])
    # packed A: rows of ABL, zero padding
    ss, ldk = n - fs, (fs + 63) // 64 * 64
    packed = lp[:((ss + 127) // 128 * 128) * ldk].reshape(-1, ldk)
    assert not packed[ss:].any() and not packed[:, fs:].any()
    if not fp16:
        got = np.ascontiguousarray(packed[:ss, :fs].astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        assert np.array_equal(got, bf16_round(ABL))
