"""uhm_amd -- MI355X multifrontal factorize / solve behind the UHM mesh API.

Python mirror of the reference's mesh interface (UHM/include/uhm/mesh/mesh.hxx:58-301,
driven as in UHM/test/performance/uhmtest.cxx:28-243) over the C ABI of
libuhm_amd.so (include/uhm_amd.h).  Every numeric operation runs in HIP kernels
on the GPU; there is no CPU fallback: calls that need the device raise when it
is absent.

    m = Mesh()
    m.build_synthetic(kind=2, dims=(16, 16, 16), p=3)     # or add_node/add_element
    m.lock()
    m.create_matrix_without_buffer(UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=2, alpha=64, beta=0.1, scale_kind=1, p0=1.0)
    m.random_matrix()
    m.set_rhs()
    m.lu_piv_with_free(); m.solve_lu_piv(); m.check_lu_piv()
    m.get_residual(), m.backward_error()
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libuhm_amd.so")

UHM_REAL = 102
UHM_CHOL = 1
UHM_LU_PIV = 3
UHM_LHS = 1
UHM_RHS = 2
UHM_PHYSICS_SINGLE = 1
UHM_PHYSICS_MULTI = 2

_lib = None

_P = C.POINTER(C.c_long)
_I = C.POINTER(C.c_int)
_D = C.POINTER(C.c_double)
_LL = C.POINTER(C.c_longlong)

# name -> argument types; every entry point returns void unless listed in _RET
_SIGS = {
    "uhm_initialize_fla_": [], "uhm_finalize_fla_": [],
    "uhm_set_num_threads_": [_I], "uhm_get_num_threads_": [_I],
    "uhm_set_hier_block_size_": [_I], "uhm_get_hier_block_size_": [_I],
    "uhm_reset_flop_": [], "uhm_get_flop_": [_D],
    "uhm_get_buffer_used_": [_D], "uhm_get_max_buffer_used_": [_D],
    "uhm_is_multithreading_enable_": [_I], "uhm_is_hier_matrix_enable_": [_I],
    "uhm_timer_": [_D],
    "uhm_mesh_create_": [_P], "uhm_mesh_delete_": [_P],
    "uhm_add_element_": [_P, _P],
    "uhm_add_node_": [_P, _I, _I, _I, _I, _P],
    "uhm_element_add_node_": [_P, _P],
    "uhm_copy_in_": [_P, _P, _I, _I, _I, _I, _I, _I, _D],
    "uhm_copy_out_": [_P, _P, _I, _I, _I, _I, _I, _I, _D],
    "uhm_mesh_disp_": [_P], "uhm_get_n_dof_": [_P, _I], "uhm_get_n_nonzero_": [_P, _I],
    "uhm_estimate_cost_": [_P, _I, _I, _I, _D, _D, C.POINTER(C.c_uint), _D],
    "uhm_mesh_import_": [_P, C.c_char_p],
    "uhm_build_tree_": [_P],
    "uhm_create_matrix_without_buffer_": [_P, _I, _I],
    "uhm_create_leaf_matrix_buffer_": [_P],
    "uhm_create_element_matrix_buffer_": [_P, _I],
    "uhm_set_rhs_": [_P], "uhm_random_spd_matrix_": [_P], "uhm_random_matrix_": [_P],
    "uhm_triangularize_": [_P], "uhm_lock_": [_P], "uhm_unlock_": [_P],
    "uhm_chol_with_free_": [_P], "uhm_lu_piv_with_free_": [_P],
    "uhm_solve_chol_": [_P], "uhm_solve_lu_piv_": [_P],
    "uhm_check_chol_": [_P], "uhm_check_lu_piv_": [_P],
    "uhm_get_residual_": [_P, _D],
    "uhm_element_set_parent_": [_P, _P], "uhm_element_add_child_": [_P, _P],
    "uhm_build_synthetic_mesh_": [_P, _I, _I, _I, _LL],
    "uhm_set_random_params_": [_P, _LL, _D, _D, _I, _D, _D],
    "uhm_get_info_": [_P, _I], "uhm_get_backward_error_": [_P, _D],
    "uhm_get_timing_": [_P, _D, _D],
    "uhm_get_last_error_": [C.c_char_p, C.c_int],
    "uhm_get_n_fronts_": [_P, _I], "uhm_get_front_info_": [_P, _I, _I],
    "uhm_get_front_nodes_": [_P, _I, _I, _I, _I, _I],
    "uhm_get_mapper_": [_P, _I, _I, _I],
    "uhm_get_node_offset_": [_P, _I, _I, _I],
    "uhm_fetch_front_": [_P, _I, _D], "uhm_fetch_pivots_": [_P, _I, _I],
    "uhm_fetch_stage_": [_P, _I, _D], "uhm_fetch_rhs_front_": [_P, _I, _D],
    "uhm_fetch_x_global_": [_P, _D], "uhm_fetch_b_global_": [_P, _D],
    "uhm_get_csr_size_": [_P, _LL], "uhm_fetch_csr_": [_P, _LL, _I, _D],
    "uhm_export_connectivity_": [_P, _I, _I, _I, _I, _I, _I, _I, _I, _I],
    "uhm_set_profile_": [_P, _I], "uhm_get_kernel_stats_": [_P, _D, _D, _D, _I],
    "uhm_set_precision_": [_P, _I, _I], "uhm_set_graph_": [_P, _I],
    "uhm_set_refinement_": [_P, _I, _I, _D], "uhm_get_refine_solves_": [_P, _I],
    "uhm_set_tuning_": [_P, _I, _I], "uhm_get_tuning_": [_P, _I, _I], "uhm_autotune_": [_P, _I, _D],
    "uhm_improve_chol_": [_P, _D, _I, _I], "uhm_improve_lu_piv_": [_P, _D, _I, _I],
    "uhm_get_refine_history_": [_P, _D, _I, _D],
    "uhm_set_node_": [_P, _I, _I, _I, _I], "uhm_backup_": [_P, _P], "uhm_check_reuse_": [_P, _P],
    "uhm_get_n_reusable_": [_P, _I], "uhm_get_n_reused_": [_P, _I],
}

KERNEL_KINDS = ("assemble", "lu_panel", "lu_rows", "chol_diag", "chol_trsm", "gemm", "schur", "lu_front",
                "gemm_diag", "chol_group", "lp_pack", "chol_small", "schur_lp", "p32_pack", "schur_p32",
                "gemm_p32")
_RET = {"uhm_get_last_error_": C.c_int}


PREC_FP64, PREC_MIXED, PREC_MIXED_BF16, PREC_MIXED_FP16 = 0, 1, 2, 3
# front size classes of the per-class schedule knobs (tuning keys 100 + c,
# 110 + c, 120 + c): fs <= 64, fs <= 512, larger
N_FRONT_CLASSES = 3
FRONT_CLASS_LIMITS = (64, 512)
REFINE_CLASSICAL, REFINE_GMRES, REFINE_AUTO = 0, 1, 2


class UHMError(RuntimeError):
    pass


def library_path():
    return LIB_PATH


def load(path=None):
    """Load libuhm_amd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise UHMError("libuhm_amd.so not built: run uhm_amd.build.build() (%s)" % path)
    lib = C.CDLL(path)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RET.get(name, None)
    _lib = lib
    return lib


def exported_symbols():
    return sorted(_SIGS)


def _err():
    buf = C.create_string_buffer(4096)
    n = _lib.uhm_get_last_error_(buf, 4096)
    if n:
        raise UHMError(buf.value.decode(errors="replace"))


def _i(v):
    return C.byref(C.c_int(int(v)))


def _call(name, *args):
    getattr(load(), name)(*args)
    _err()


def _dptr(a):
    return a.ctypes.data_as(_D)


def _iptr(a):
    return a.ctypes.data_as(_I)


class Node:
    def __init__(self, handle, id, phy, n_dof):
        self.handle, self.id, self.phy, self.n_dof = handle, id, phy, n_dof


class Element:
    def __init__(self, handle, index):
        self.handle, self.index = handle, index
        self.nodes = []

    def add_node(self, node):
        _call("uhm_element_add_node_", C.byref(self.handle), C.byref(node.handle))
        self.nodes.append(node)

    def set_parent(self, parent):
        _call("uhm_element_set_parent_", C.byref(self.handle), C.byref(parent.handle))

    def add_child(self, child):
        _call("uhm_element_add_child_", C.byref(self.handle), C.byref(child.handle))


class Mesh:
    """The reference's Mesh_ call sequence over the device solver."""

    def __init__(self):
        load()
        self.handle = C.c_long(0)
        _call("uhm_mesh_create_", C.byref(self.handle))
        self.elements = []
        self.n_rhs = 1

    def __del__(self):
        try:
            if _lib is not None and self.handle.value:
                _lib.uhm_mesh_delete_(C.byref(self.handle))
                _lib.uhm_get_last_error_(None, 0)
        except Exception:
            pass

    # ---- construction
    def add_node(self, id, n_dof, p=None, phy=0):
        h = C.c_long(0)
        _call("uhm_add_node_", C.byref(self.handle), _i(id), _i(phy), _i(n_dof),
              _i(n_dof if p is None else p), C.byref(h))
        return Node(h, id, phy, n_dof)

    def add_element(self):
        h = C.c_long(0)
        _call("uhm_add_element_", C.byref(self.handle), C.byref(h))
        e = Element(h, len(self.elements))
        self.elements.append(e)
        return e

    def build_synthetic(self, kind, dims, p, seed=0):
        d = np.zeros(3, dtype=np.int32)
        d[:len(dims)] = dims
        _call("uhm_build_synthetic_mesh_", C.byref(self.handle), _i(kind), _iptr(d), _i(p),
              C.byref(C.c_longlong(seed)))

    def import_file(self, path):
        _call("uhm_mesh_import_", C.byref(self.handle), path.encode())

    def build_tree(self):
        _call("uhm_build_tree_", C.byref(self.handle))

    def set_node(self, id, n_dof, p=None, phy=0):
        """New dof count (and order) of an existing node; the mesh must be
        unlocked."""
        _call("uhm_set_node_", C.byref(self.handle), _i(id), _i(phy), _i(n_dof), _i(n_dof if p is None else p))

    def backup(self, other):
        """Keep this mesh's connectivity in `other` (Mesh_::backup)."""
        _call("uhm_backup_", C.byref(self.handle), C.byref(other.handle))

    def check_reuse(self, other):
        """Flag the elements whose factored fronts the next factorization
        keeps (Mesh_::check_reuse against the connectivity `other` keeps)."""
        _call("uhm_check_reuse_", C.byref(self.handle), C.byref(other.handle))

    def n_reusable(self):
        v = C.c_int(0)
        _call("uhm_get_n_reusable_", C.byref(self.handle), C.byref(v))
        return v.value

    def n_reused(self):
        v = C.c_int(0)
        _call("uhm_get_n_reused_", C.byref(self.handle), C.byref(v))
        return v.value

    def lock(self):
        _call("uhm_lock_", C.byref(self.handle))

    def unlock(self):
        _call("uhm_unlock_", C.byref(self.handle))

    # ---- fronts and data
    def create_matrix_without_buffer(self, datatype=UHM_REAL, n_rhs=1):
        self.n_rhs = n_rhs
        _call("uhm_create_matrix_without_buffer_", C.byref(self.handle), _i(datatype), _i(n_rhs))

    def create_matrix_buffer(self, is_schur=0):
        _call("uhm_create_element_matrix_buffer_", C.byref(self.handle), _i(is_schur))

    def set_random_params(self, seed=1, alpha=0.0, beta=0.0, scale_kind=0, p0=0.0, p1=0.0):
        _call("uhm_set_random_params_", C.byref(self.handle), C.byref(C.c_longlong(seed)),
              C.byref(C.c_double(alpha)), C.byref(C.c_double(beta)), _i(scale_kind),
              C.byref(C.c_double(p0)), C.byref(C.c_double(p1)))

    def random_matrix(self):
        _call("uhm_random_matrix_", C.byref(self.handle))

    def random_spd_matrix(self):
        _call("uhm_random_spd_matrix_", C.byref(self.handle))

    def copy_in(self, elt, nods, side, buf, node_disp=UHM_PHYSICS_SINGLE):
        buf = np.asfortranarray(buf, dtype=np.float64)
        if buf.ndim == 1:
            buf = buf.reshape(-1, 1, order="F")
        nods = np.ascontiguousarray(nods, dtype=np.int32)
        _call("uhm_copy_in_", C.byref(self.handle), C.byref(elt.handle), _i(UHM_REAL),
              _i(buf.shape[0]), _i(buf.shape[1]), _i(node_disp), _iptr(nods), _i(side), _dptr(buf))

    def copy_out(self, elt, nods, side, m, n, node_disp=UHM_PHYSICS_SINGLE):
        out = np.zeros((m, n), dtype=np.float64, order="F")
        nods = np.ascontiguousarray(nods, dtype=np.int32)
        _call("uhm_copy_out_", C.byref(self.handle), C.byref(elt.handle), _i(UHM_REAL),
              _i(m), _i(n), _i(node_disp), _iptr(nods), _i(side), _dptr(out))
        return out

    def set_rhs(self):
        _call("uhm_set_rhs_", C.byref(self.handle))

    # ---- factorize / solve / check
    def lu_piv_with_free(self):
        _call("uhm_lu_piv_with_free_", C.byref(self.handle))
        return self.info()

    def chol_with_free(self):
        _call("uhm_chol_with_free_", C.byref(self.handle))
        return self.info()

    def solve_lu_piv(self):
        _call("uhm_solve_lu_piv_", C.byref(self.handle))

    def solve_chol(self):
        _call("uhm_solve_chol_", C.byref(self.handle))

    def check_lu_piv(self):
        _call("uhm_check_lu_piv_", C.byref(self.handle))

    def check_chol(self):
        _call("uhm_check_chol_", C.byref(self.handle))

    def get_residual(self):
        r = C.c_double(0)
        _call("uhm_get_residual_", C.byref(self.handle), C.byref(r))
        return r.value

    def backward_error(self):
        r = C.c_double(0)
        _call("uhm_get_backward_error_", C.byref(self.handle), C.byref(r))
        return r.value

    def info(self):
        r = C.c_int(0)
        _call("uhm_get_info_", C.byref(self.handle), C.byref(r))
        return r.value

    def timing(self):
        a, b = C.c_double(0), C.c_double(0)
        _call("uhm_get_timing_", C.byref(self.handle), C.byref(a), C.byref(b))
        return a.value, b.value

    def set_precision(self, mode, top_levels=2):
        """mode PREC_FP64 (0), PREC_MIXED (1: FP32 fronts below the top
        `top_levels` levels, FP32 MFMA Schur updates), PREC_MIXED_BF16 (2:
        as 1 with BF16-operand, FP32-accumulate Schur complement products) or
        PREC_MIXED_FP16 (3: as 2 with FP16 operands, scaled per front)."""
        _call("uhm_set_precision_", C.byref(self.handle), _i(mode), _i(top_levels))

    def improve(self, method, tol=1e-12, max_iter=20):
        """Iterative refinement to backward error tol; returns the number of
        corrections applied."""
        n = C.c_int(0)
        name = "uhm_improve_chol_" if method == UHM_CHOL else "uhm_improve_lu_piv_"
        _call(name, C.byref(self.handle), C.byref(C.c_double(tol)), _i(max_iter), C.byref(n))
        return n.value

    def refine_history(self):
        """(backward error before each correction and at the end, ms)."""
        eta = np.zeros(256)
        n, ms = C.c_int(len(eta)), C.c_double(0)
        _call("uhm_get_refine_history_", C.byref(self.handle), _dptr(eta), C.byref(n), C.byref(ms))
        return eta[:min(n.value, len(eta))].copy(), ms.value

    def set_refinement(self, kind, restart=20, inner_tol=1e-6):
        """Correction kind of improve(): REFINE_CLASSICAL, REFINE_GMRES
        (GMRES-IR, restart length, inner tolerance) or REFINE_AUTO."""
        _call("uhm_set_refinement_", C.byref(self.handle), _i(kind), _i(restart), C.byref(C.c_double(inner_tol)))

    def refine_solves(self):
        """Level-solve applications of M^-1 in the last improve()."""
        n = C.c_int(0)
        _call("uhm_get_refine_solves_", C.byref(self.handle), C.byref(n))
        return n.value

    def set_tuning(self, key, value):
        _call("uhm_set_tuning_", C.byref(self.handle), _i(key), _i(value))

    def get_tuning(self, key):
        v = C.c_int(0)
        _call("uhm_get_tuning_", C.byref(self.handle), _i(key), C.byref(v))
        return v.value

    def autotune(self, method):
        """Times factorizations over the schedule knobs, keeps the fastest;
        returns (best ms, {knob: value})."""
        t = C.c_double(0)
        _call("uhm_autotune_", C.byref(self.handle), _i(method), C.byref(t))
        knobs = {"chol_group": self.get_tuning(1), "tile_threshold": self.get_tuning(2),
                 "lookahead": self.get_tuning(5)}
        for c in range(N_FRONT_CLASSES):
            knobs["class%d" % c] = {"chol_group": self.get_tuning(100 + c), "tile_threshold": self.get_tuning(110 + c),
                                    "group_launch": self.get_tuning(120 + c)}
        return t.value, knobs

    def set_graph(self, on=True):
        """Capture the factorization / solves as HIP graphs and replay them."""
        _call("uhm_set_graph_", C.byref(self.handle), _i(1 if on else 0))

    def set_profile(self, on=True):
        _call("uhm_set_profile_", C.byref(self.handle), _i(1 if on else 0))

    def kernel_stats(self):
        """{kind: (ms, algorithmic flops, algorithmic bytes, launches)} of the
        last profiled factorization."""
        ms, fl, by = (np.zeros(len(KERNEL_KINDS)) for _ in range(3))
        ln = np.zeros(len(KERNEL_KINDS), np.int32)
        _call("uhm_get_kernel_stats_", C.byref(self.handle), _dptr(ms), _dptr(fl), _dptr(by), _iptr(ln))
        return {k: (float(ms[i]), float(fl[i]), float(by[i]), int(ln[i])) for i, k in enumerate(KERNEL_KINDS)}

    # ---- queries
    def get_n_dof(self):
        r = C.c_int(0)
        _call("uhm_get_n_dof_", C.byref(self.handle), C.byref(r))
        return r.value

    def estimate_cost(self, method, n_rhs=1):
        fd, fsv, buf = C.c_double(0), C.c_double(0), C.c_double(0)
        nnz = C.c_uint(0)
        _call("uhm_estimate_cost_", C.byref(self.handle), _i(method), _i(UHM_REAL), _i(n_rhs),
              C.byref(fd), C.byref(fsv), C.byref(nnz), C.byref(buf))
        return fd.value, fsv.value, nnz.value, buf.value

    def n_fronts(self):
        r = C.c_int(0)
        _call("uhm_get_n_fronts_", C.byref(self.handle), C.byref(r))
        return r.value

    def front_info(self, f):
        out = np.zeros(9, dtype=np.int32)
        _call("uhm_get_front_info_", C.byref(self.handle), _i(f), _iptr(out))
        keys = ("id", "fs", "ss", "level", "parent", "generation", "n_factor", "n_schur", "leaf")
        return dict(zip(keys, (int(v) for v in out)))

    def fronts(self):
        return [self.front_info(f) for f in range(self.n_fronts())]

    def front_nodes(self, f, which):
        n = C.c_int(0)
        _call("uhm_get_front_nodes_", C.byref(self.handle), _i(f), _i(which), None, None, C.byref(n))
        ids = np.zeros(max(1, n.value), dtype=np.int32)
        offs = np.zeros(max(1, n.value), dtype=np.int32)
        _call("uhm_get_front_nodes_", C.byref(self.handle), _i(f), _i(which), _iptr(ids), _iptr(offs),
              C.byref(n))
        return ids[:n.value].copy(), offs[:n.value].copy()

    def mapper(self, child_front):
        n = C.c_int(0)
        _call("uhm_get_mapper_", C.byref(self.handle), _i(child_front), None, C.byref(n))
        runs = np.zeros(max(4, 4 * n.value), dtype=np.int32)
        _call("uhm_get_mapper_", C.byref(self.handle), _i(child_front), _iptr(runs), C.byref(n))
        return runs[:4 * n.value].reshape(-1, 4).copy()

    def node_offset(self, id, phy=0):
        r = C.c_int(0)
        _call("uhm_get_node_offset_", C.byref(self.handle), _i(id), _i(phy), C.byref(r))
        return r.value

    def fetch_front(self, f):
        n = self.front_info(f)
        n = n["fs"] + n["ss"]
        out = np.zeros((n, n), dtype=np.float64, order="F")
        _call("uhm_fetch_front_", C.byref(self.handle), _i(f), _dptr(out))
        return out

    def fetch_stage(self, f):
        n = self.front_info(f)
        n = n["fs"] + n["ss"]
        out = np.zeros((n, n), dtype=np.float64, order="F")
        _call("uhm_fetch_stage_", C.byref(self.handle), _i(f), _dptr(out))
        return out

    def fetch_rhs_front(self, f):
        n = self.front_info(f)
        n = n["fs"] + n["ss"]
        out = np.zeros((n, self.n_rhs), dtype=np.float64, order="F")
        _call("uhm_fetch_rhs_front_", C.byref(self.handle), _i(f), _dptr(out))
        return out

    def fetch_pivots(self, f):
        fs = self.front_info(f)["fs"]
        out = np.zeros(max(1, fs), dtype=np.int32)
        _call("uhm_fetch_pivots_", C.byref(self.handle), _i(f), _iptr(out))
        return out[:fs].copy()

    def fetch_x(self):
        n = self.get_n_dof()
        out = np.zeros((n, self.n_rhs), dtype=np.float64, order="F")
        _call("uhm_fetch_x_global_", C.byref(self.handle), _dptr(out))
        return out

    def fetch_b(self):
        n = self.get_n_dof()
        out = np.zeros((n, self.n_rhs), dtype=np.float64, order="F")
        _call("uhm_fetch_b_global_", C.byref(self.handle), _dptr(out))
        return out

    def export_connectivity(self):
        """Nodes (creation order), leaf node lists, parents and children."""
        sz = np.zeros(4, dtype=np.int32)
        h = C.byref(self.handle)
        _call("uhm_export_connectivity_", h, _iptr(sz), None, None, None, None, None, None, None, None)
        nn, ne, ln, lc = (int(v) for v in sz)
        ids, ndof, kind = (np.zeros(max(1, nn), np.int32) for _ in range(3))
        eptr, cptr = np.zeros(ne + 1, np.int32), np.zeros(ne + 1, np.int32)
        enode, cidx = np.zeros(max(1, ln), np.int32), np.zeros(max(1, lc), np.int32)
        parent = np.zeros(max(1, ne), np.int32)
        _call("uhm_export_connectivity_", h, _iptr(sz), _iptr(ids), _iptr(ndof), _iptr(kind), _iptr(eptr),
              _iptr(enode), _iptr(parent), _iptr(cptr), _iptr(cidx))
        return dict(node_id=ids[:nn], ndof=ndof[:nn], kind=kind[:nn], eptr=eptr, enode=enode[:ln],
                    parent=parent[:ne], cptr=cptr, cidx=cidx[:lc])

    def fetch_csr(self):
        nnz = C.c_longlong(0)
        _call("uhm_get_csr_size_", C.byref(self.handle), C.byref(nnz))
        n = self.get_n_dof()
        rp = np.zeros(n + 1, dtype=np.int64)
        col = np.zeros(max(1, nnz.value), dtype=np.int32)
        val = np.zeros(max(1, nnz.value), dtype=np.float64)
        _call("uhm_fetch_csr_", C.byref(self.handle), rp.ctypes.data_as(_LL), _iptr(col), _dptr(val))
        return rp, col[:nnz.value], val[:nnz.value]


def set_hier_block_size(b):
    _call("uhm_set_hier_block_size_", _i(b))


def get_hier_block_size():
    r = C.c_int(0)
    _call("uhm_get_hier_block_size_", C.byref(r))
    return r.value


def buffer_used():
    """Device bytes the library holds now (every mesh of the process)."""
    r = C.c_double(0)
    _call("uhm_get_buffer_used_", C.byref(r))
    return r.value


def max_buffer_used():
    """High-water mark of buffer_used()."""
    r = C.c_double(0)
    _call("uhm_get_max_buffer_used_", C.byref(r))
    return r.value
