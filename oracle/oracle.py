"""ctypes front-end of the CPU oracle (oracle/uhm_oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle restates the reference's multifrontal
algorithm on the CPU (see the header of uhm_oracle.c for the file:line of
every step).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline use it, always as the checker or the timed baseline, never as the
product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_I = C.POINTER(C.c_int)
_D = C.POINTER(C.c_double)
_lib = None


def build():
    # -B: always recompile (the caller asked for a build)
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    lib.orc_analyze.restype = C.c_void_p
    lib.orc_analyze.argtypes = [C.c_int, _I, _I, C.c_int, _I, _I, _I, _I, _I]
    lib.orc_free.argtypes = [C.c_void_p]
    for name in ("orc_n_levels", "orc_max_threads"):
        getattr(lib, name).restype = C.c_int
    lib.orc_n_levels.argtypes = [C.c_void_p]
    lib.orc_level.restype = C.c_int
    lib.orc_level.argtypes = [C.c_void_p, C.c_int, _I]
    lib.orc_element.argtypes = [C.c_void_p, C.c_int, _I]
    lib.orc_lists.argtypes = [C.c_void_p, C.c_int, _I, _I, _I, _I]
    lib.orc_node_offset.restype = C.c_int
    lib.orc_node_offset.argtypes = [C.c_void_p, C.c_int]
    lib.orc_mapper.restype = C.c_int
    lib.orc_mapper.argtypes = [C.c_void_p, C.c_int, _I]
    lib.orc_init_numeric.argtypes = [C.c_void_p, C.c_int]
    lib.orc_copy_in.restype = C.c_int
    lib.orc_copy_in.argtypes = [C.c_void_p, C.c_int, C.c_int, _I, C.c_int, C.c_int, _D]
    lib.orc_set_front.argtypes = [C.c_void_p, C.c_int, _D]
    lib.orc_set_rhs_front.argtypes = [C.c_void_p, C.c_int, _D]
    lib.orc_factor.restype = C.c_int
    lib.orc_factor.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.orc_factor_levels.restype = C.c_int
    lib.orc_factor_levels.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.orc_level_flops.restype = C.c_double
    lib.orc_level_flops.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    lib.orc_solve.argtypes = [C.c_void_p, C.c_int]
    lib.orc_get_front.argtypes = [C.c_void_p, C.c_int, _D]
    lib.orc_get_pivots.argtypes = [C.c_void_p, C.c_int, _I]
    lib.orc_get_margins.argtypes = [C.c_void_p, C.c_int, _D]
    lib.orc_get_input.argtypes = [C.c_void_p, C.c_int, _D]
    lib.orc_assemble.argtypes = [C.c_void_p, C.c_int, _D]
    lib.orc_set_factored.argtypes = [C.c_void_p, C.c_int, _D]
    lib.orc_get_x_global.argtypes = [C.c_void_p, _D, C.c_int]
    lib.orc_flops.restype = C.c_double
    lib.orc_flops.argtypes = [C.c_void_p, C.c_int]
    lib.orc_set_threads.argtypes = [C.c_int]
    lib.orc_dense_lu.restype = C.c_int
    lib.orc_dense_lu.argtypes = [_D, C.c_int, C.c_int, C.c_int, _I]
    lib.orc_panel_lu.restype = C.c_int
    lib.orc_panel_lu.argtypes = [_D, C.c_int, C.c_int, C.c_int, _I, _D]
    lib.orc_dense_chol.restype = C.c_int
    lib.orc_dense_chol.argtypes = [_D, C.c_int, C.c_int, C.c_int]
    _lib = lib
    return lib


def _ip(a):
    return a.ctypes.data_as(_I)


def _dp(a):
    return a.ctypes.data_as(_D)


def set_threads(n):
    load().orc_set_threads(int(n))


def max_threads():
    return load().orc_max_threads()


def dense_lu(F, fs, b=256):
    """Front LU (lu_piv_hier semantics) of a copy of F; returns (F, p, info)."""
    A = np.array(F, dtype=np.float64, order="F", copy=True)
    n = A.shape[0]
    p = np.zeros(max(1, fs), dtype=np.int32)
    info = load().orc_dense_lu(_dp(A), n, fs, b, _ip(p))
    return A, p[:fs], info


def panel_lu(A, kb):
    """Unblocked GEPP panel (getf2) of a copy of the m x kb block A;
    returns (A, relative pivots, pivot margins, info)."""
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m = A.shape[0]
    p = np.zeros(max(1, kb), dtype=np.int32)
    marg = np.zeros(max(1, kb), dtype=np.float64)
    info = load().orc_panel_lu(_dp(A), m, m, kb, _ip(p), _dp(marg))
    return A, p[:kb], marg[:kb], info


def dense_chol(F, fs, b=256):
    A = np.array(F, dtype=np.float64, order="F", copy=True)
    info = load().orc_dense_chol(_dp(A), A.shape[0], fs, b)
    return A, info


class OracleMesh:
    """Analysis + numeric state of one mesh in the oracle."""

    def __init__(self, ndof, kind, eptr, enode, parent, cptr, cidx):
        lib = load()
        self.ndof = np.ascontiguousarray(ndof, np.int32)
        self.kind = np.ascontiguousarray(kind, np.int32)
        self.eptr = np.ascontiguousarray(eptr, np.int32)
        self.enode = np.ascontiguousarray(enode, np.int32)
        self.parent = np.ascontiguousarray(parent, np.int32)
        self.cptr = np.ascontiguousarray(cptr, np.int32)
        self.cidx = np.ascontiguousarray(cidx, np.int32)
        self.n_elem = len(self.parent)
        self.h = lib.orc_analyze(len(self.ndof), _ip(self.ndof), _ip(self.kind), self.n_elem, _ip(self.eptr),
                                 _ip(self.enode), _ip(self.parent), _ip(self.cptr), _ip(self.cidx))
        self.nrhs = None

    @classmethod
    def from_connectivity(cls, c):
        return cls(c["ndof"], c["kind"], c["eptr"], c["enode"], c["parent"], c["cptr"], c["cidx"])

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_free(self.h)
            self.h = None

    # ---- analysis
    def levels(self):
        out = []
        for L in range(_lib.orc_n_levels(self.h)):
            n = _lib.orc_level(self.h, L, None)
            a = np.zeros(max(1, n), np.int32)
            _lib.orc_level(self.h, L, _ip(a))
            out.append(a[:n].copy())
        return out

    def element(self, e):
        out = np.zeros(5, np.int32)
        _lib.orc_element(self.h, e, _ip(out))
        gen, fs, ss, nf, ns = (int(v) for v in out)
        fac, facoff = np.zeros(max(1, nf), np.int32), np.zeros(max(1, nf), np.int32)
        sch, schoff = np.zeros(max(1, ns), np.int32), np.zeros(max(1, ns), np.int32)
        _lib.orc_lists(self.h, e, _ip(fac), _ip(facoff), _ip(sch), _ip(schoff))
        return dict(generation=gen, fs=fs, ss=ss, factor=fac[:nf].copy(), factor_off=facoff[:nf].copy(),
                    schur=sch[:ns].copy(), schur_off=schoff[:ns].copy())

    def node_offset(self, node):
        return _lib.orc_node_offset(self.h, int(node))

    def mapper(self, child):
        e = self.element(child)
        runs = np.zeros(4 * max(1, len(e["schur"])), np.int32)
        n = _lib.orc_mapper(self.h, child, _ip(runs))
        return runs[:4 * n].reshape(-1, 4).copy()

    def flops(self, method):
        return _lib.orc_flops(self.h, method)

    # ---- numerics
    def init_numeric(self, nrhs=1):
        self.nrhs = nrhs
        _lib.orc_init_numeric(self.h, nrhs)

    def copy_in(self, e, nods, side, buf):
        buf = np.asfortranarray(buf, dtype=np.float64)
        nods = np.ascontiguousarray(nods, np.int32)
        r = _lib.orc_copy_in(self.h, e, len(nods), _ip(nods), side, buf.shape[0], _dp(buf))
        if r:
            raise RuntimeError("oracle copy_in: node list does not match the element")

    def set_front(self, e, F):
        F = np.asfortranarray(F, dtype=np.float64)
        _lib.orc_set_front(self.h, e, _dp(F))

    def set_rhs_front(self, e, b):
        b = np.asfortranarray(b, dtype=np.float64)
        _lib.orc_set_rhs_front(self.h, e, _dp(b))

    def factor(self, method, b=256):
        return _lib.orc_factor(self.h, method, b)

    def factor_levels(self, method, b, L0, L1):
        return _lib.orc_factor_levels(self.h, method, b, L0, L1)

    def level_flops(self, method, L0, L1):
        return _lib.orc_level_flops(self.h, method, L0, L1)

    def solve(self, method):
        _lib.orc_solve(self.h, method)

    def front(self, e):
        el = self.element(e)
        n = el["fs"] + el["ss"]
        out = np.zeros((n, n), np.float64, order="F")
        _lib.orc_get_front(self.h, e, _dp(out))
        return out

    def pivots(self, e):
        fs = self.element(e)["fs"]
        out = np.zeros(max(1, fs), np.int32)
        _lib.orc_get_pivots(self.h, e, _ip(out))
        return out[:fs].copy()

    def margins(self, e):
        """Relative margin (|best| - |next|) / |best| of every pivot choice
        of the last LU factorization of front e."""
        fs = self.element(e)["fs"]
        out = np.zeros(max(1, fs), np.float64)
        _lib.orc_get_margins(self.h, e, _dp(out))
        return out[:fs].copy()

    def _square(self, e):
        el = self.element(e)
        n = el["fs"] + el["ss"]
        return np.zeros((n, n), np.float64, order="F")

    def input(self, e):
        """Assembly input of front e (its copy-in, front order)."""
        out = self._square(e)
        _lib.orc_get_input(self.h, e, _dp(out))
        return out

    def set_factored(self, e, F):
        """Working values of front e (e.g. factored on the GPU), read by
        assemble() of its parent."""
        F = np.asfortranarray(F, dtype=np.float64)
        _lib.orc_set_factored(self.h, e, _dp(F))

    def assemble(self, e):
        """Extend-add of front e (merge_A: 0 + own input + children's ABR in
        children order), without factoring it; its children must have been
        factored."""
        out = self._square(e)
        _lib.orc_assemble(self.h, e, _dp(out))
        return out

    def x_global(self, ndof):
        out = np.zeros((ndof, self.nrhs), np.float64, order="F")
        _lib.orc_get_x_global(self.h, _dp(out), ndof)
        return out
