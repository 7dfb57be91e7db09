"""Host restatement of the synthetic element generator (test infrastructure).

The device generator draws every element matrix from a Philox-4x32-10
counter stream (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
easy as 1, 2, 3", SC'11): counter (draw index, element id, stream, kind),
key (seed low, seed high); Box-Muller normals from two 32-bit words;
K_e = s_e (M^T M + shift I + beta S), b_e ~ N(0, 1).  This module restates
that stream in numpy so that tests can pin the device's draws and the
matrices built from them.  Streams: 0 the entries of M, 1 those of S, 2 the
element scale, 3 the right-hand side.
"""
import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_TWO_M32 = 2.3283064365386963e-10  # 2^-32


def philox4x32_10(c, k):
    """Philox-4x32-10 of counters c = (c0, c1, c2, c3) (uint32 arrays or
    scalars) under key k = (k0, k1); returns the four output words."""
    x, y, z, w = (np.asarray(v, np.uint64) & np.uint64(0xFFFFFFFF) for v in c)
    k0, k1 = int(k[0]) & 0xFFFFFFFF, int(k[1]) & 0xFFFFFFFF
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = _M0 * x
        p1 = _M1 * z
        hi0, lo0 = p0 >> np.uint64(32), p0 & mask
        hi1, lo1 = p1 >> np.uint64(32), p1 & mask
        x, y, z, w = hi1 ^ y ^ np.uint64(k0), lo1, hi0 ^ w ^ np.uint64(k1), lo0
        k0 = (k0 + _W0) & 0xFFFFFFFF
        k1 = (k1 + _W1) & 0xFFFFFFFF
    return tuple(v.astype(np.uint32) for v in (x, y, z, w))


def _key(seed):
    return (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)


def normals(seed, elem, stream, idx):
    """N(0, 1) draws idx (array) of one element and stream: counter
    (idx >> 1, elem, stream, 0), Box-Muller on words x and y (cosine for even
    idx, sine for odd)."""
    idx = np.asarray(idx, np.uint64)
    n = idx.shape
    r = philox4x32_10((idx >> np.uint64(1), np.full(n, elem), np.full(n, stream), np.zeros(n)), _key(seed))
    u1 = (r[0].astype(np.float64) + 1.0) * _TWO_M32
    u2 = r[1].astype(np.float64) * _TWO_M32
    rad = np.sqrt(-2.0 * np.log(u1))
    th = 6.283185307179586 * u2
    return np.where(idx & np.uint64(1), rad * np.sin(th), rad * np.cos(th))


def uniform(seed, elem, stream, idx):
    """U(0, 1) draw: counter (idx, elem, stream, 1), word x."""
    r = philox4x32_10((idx, elem, stream, 1), _key(seed))
    return (float(r[0]) + 0.5) * _TWO_M32


def element_scale(seed, elem, scale_kind, p0, p1):
    if scale_kind == 1:
        return float(np.exp(p0 * normals(seed, elem, 2, [0])[0]))
    if scale_kind == 2:
        return float(10.0 ** (p0 + (p1 - p0) * uniform(seed, elem, 2, 0)))
    return 1.0


def element_matrix(seed, elem, m, alpha, beta, scale_kind, p0, p1):
    """K_e = s_e (M^T M + shift I + beta S) in the element's own dof order;
    M(i, j) is draw i + j m of stream 0, S(i, j) of stream 1; shift = alpha,
    or |alpha| m for alpha < 0."""
    idx = np.arange(m * m)
    M = normals(seed, elem, 0, idx).reshape(m, m, order="F")
    K = M.T @ M
    K += (-alpha * m if alpha < 0 else alpha) * np.eye(m)
    if beta:
        K += beta * normals(seed, elem, 1, idx).reshape(m, m, order="F")
    return element_scale(seed, elem, scale_kind, p0, p1) * K


def element_rhs(seed, elem, m, nrhs):
    """b_e(i, r) = draw i + r m of stream 3."""
    return normals(seed, elem, 3, np.arange(m * nrhs)).reshape(m, nrhs, order="F")
