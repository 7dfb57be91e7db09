"""Write the golden fixtures under tests/golden/.

The reference ships no golden vectors, fixtures or known-answer tests (its
tests compare LINAL with libflame on fresh random input, and its .uhm meshes
are not distributed).  These fixtures pin the oracle with what the
mathematics fixes, computed here by independent library routines:

* lu_fronts.npz   front LU with pivoting restricted to ATL
                  (lu_piv_hier, UHM/src/matrix/uhm/fla/lu_piv/decompose.cxx:124-153):
                  scipy.linalg.lu_factor (LAPACK getrf, first maximal |a|)
                  on ATL, then ATR <- L^-1 P ATR, ABL <- ABL U^-1,
                  ABR <- ABR - ABL ATR with scipy.linalg.solve_triangular.
                  Pivots in libflame's relative format p[i] = piv[i] - i.
* chol_fronts.npz front Cholesky (chol_hier, chol/decompose.cxx:85-98):
                  numpy.linalg.cholesky on ATL, ABL <- ABL L^-T,
                  ABR <- ABR - ABL ABL^T (lower triangle).
* q2_mesh.npz     2D Q2 meshes (4x4 and 16x16 elements, the latter the
                  1089-dof CPU configuration) with seeded element matrices
                  and right-hand sides; the solution of the assembled system
                  (IJ sum of element matrices, UHM/src/interf/sparse.cxx:196-267)
                  by numpy.linalg.solve, stored per node id.
* ex1_maps.json   the two-element mesh of UHM/example/cxx/ex1.cxx:25-35,
                  its lock results derived by hand from the rules of
                  UHM/src/mesh/element.cxx:272-392 (see the derivation below).

Run:  python tests/golden/make_golden.py
"""
import json
import os

import numpy as np
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))


def front_lu_reference(F, fs):
    F = np.array(F, dtype=np.float64)
    n = F.shape[0]
    ATL, ATR = F[:fs, :fs], F[:fs, fs:]
    ABL, ABR = F[fs:, :fs], F[fs:, fs:]
    lu, piv = sla.lu_factor(ATL)
    L = np.tril(lu, -1) + np.eye(fs)
    U = np.triu(lu)
    P_ATR = ATR.copy()
    for i, pr in enumerate(piv):
        P_ATR[[i, pr]] = P_ATR[[pr, i]]
    ATR_n = sla.solve_triangular(L, P_ATR, lower=True, unit_diagonal=True)
    ABL_n = sla.solve_triangular(U.T, ABL.T, lower=True).T
    ABR_n = ABR - ABL_n @ ATR_n
    out = np.zeros((n, n))
    out[:fs, :fs] = lu
    out[:fs, fs:] = ATR_n
    out[fs:, :fs] = ABL_n
    out[fs:, fs:] = ABR_n
    return out, (piv - np.arange(fs)).astype(np.int32)


def front_chol_reference(F, fs):
    F = np.array(F, dtype=np.float64)
    n = F.shape[0]
    L = np.linalg.cholesky(F[:fs, :fs])
    ABL_n = sla.solve_triangular(L, F[fs:, :fs].T, lower=True).T
    ABR_n = F[fs:, fs:] - ABL_n @ ABL_n.T
    out = np.zeros((n, n))
    out[:fs, :fs] = L
    out[fs:, :fs] = ABL_n
    out[fs:, fs:] = np.tril(ABR_n)
    return out


def make_fronts():
    rng = np.random.default_rng(20240611)
    cases = [(1, 0), (1, 1), (5, 3), (7, 7), (9, 4), (16, 10), (33, 20), (64, 64), (70, 64), (97, 65),
             (130, 70), (150, 129)]
    lu_in, lu_out, lu_p, lu_fs, ch_in, ch_out, ch_fs = [], [], [], [], [], [], []
    for n, fs in cases:
        A = rng.standard_normal((n, n))
        o, p = front_lu_reference(A, fs)
        lu_in.append(A.ravel(order="F"))
        lu_out.append(o.ravel(order="F"))
        lu_p.append(p)
        lu_fs.append((n, fs))
        M = rng.standard_normal((n, n))
        S = M.T @ M + n * np.eye(n)
        ch_in.append(S.ravel(order="F"))
        ch_out.append(front_chol_reference(S, fs).ravel(order="F"))
        ch_fs.append((n, fs))
    np.savez_compressed(os.path.join(HERE, "lu_fronts.npz"), sizes=np.array(lu_fs, np.int32),
                        fin=np.concatenate(lu_in), fout=np.concatenate(lu_out),
                        piv=np.concatenate([p for p in lu_p if len(p)] or [np.zeros(0, np.int32)]))
    np.savez_compressed(os.path.join(HERE, "chol_fronts.npz"), sizes=np.array(ch_fs, np.int32),
                        fin=np.concatenate(ch_in), fout=np.concatenate(ch_out))


def q2_mesh(N, p=2):
    """Structured Q_p mesh on N x N elements: node id = grid index (x
    fastest), element node list in tensor order (x fastest)."""
    X = N * p + 1
    elems = []
    for ey in range(N):
        for ex in range(N):
            elems.append([(ex * p + a) + X * (ey * p + b) for b in range(p + 1) for a in range(p + 1)])
    return X * X, np.array(elems, np.int32)


def make_q2():
    rng = np.random.default_rng(7)
    out = {}
    for N, spd in ((4, False), (16, True)):
        nn, el = q2_mesh(N)
        m = el.shape[1]
        K = np.zeros((len(el), m, m))
        for e in range(len(el)):
            Mm = rng.standard_normal((m, m))
            K[e] = Mm.T @ Mm + 9.0 * np.eye(m)
            if not spd:
                K[e] += 0.1 * rng.standard_normal((m, m))
        b = rng.standard_normal((len(el), m))
        A = np.zeros((nn, nn))
        rhs = np.zeros(nn)
        for e in range(len(el)):
            A[np.ix_(el[e], el[e])] += K[e]
            rhs[el[e]] += b[e]
        x = np.linalg.solve(A, rhs)
        out["N%d_elems" % N] = el
        out["N%d_K" % N] = K
        out["N%d_b" % N] = b
        out["N%d_x" % N] = x
        out["N%d_spd" % N] = np.array(spd)
    np.savez_compressed(os.path.join(HERE, "q2_mesh.npz"), **out)


def make_ex1():
    # ex1.cxx:25-35: element 0 nodes (3,4,5 | 7,9,8 | 1), element 1 nodes
    # (4,6,5 | 10,11,9 | 2), n_dof 1,1,1 | 3,3,3 | 3.  Nodes are created in
    # the order they are first added: 3,4,5,7,9,8,1,6,10,11,2.  A parent
    # element 2 holds children [0, 1].
    #
    # Separation: nodes 4, 5, 9 are owned by both leaves -> schur in 0 and 1;
    # the rest are factor in their leaf.  The parent receives {4, 5, 9}, owned
    # by it alone -> factor there.
    # Numbering, generation 0 first, element id order, node creation order:
    #   e0: 3->0, 7->1 (3 dofs), 8->4, 1->7          (g = 10)
    #   e1: 6->10, 10->11, 11->14, 2->17             (g = 20)
    #   e2: 4->20, 5->21, 9->22                       (g = 25)
    # Arrangement (sorted by offset, cumulative local offsets):
    #   e0 factor 3:0 7:1 8:4 1:7 (fs 10), schur 4:0 5:1 9:2 (ss 5)
    #   e1 factor 6:0 10:1 11:4 2:7 (fs 10), schur 4:0 5:1 9:2 (ss 5)
    #   e2 factor 4:0 5:1 9:2 (fs 5), schur none
    # Mapper (parent 2, child 0 or 1): the three child schur nodes land at
    # parent factor offsets 0, 1, 2 with n_dof 1, 1, 3 -> one run
    # {offs_c 0, offs_p 0, fs_p 0, n_dof 5}.
    # copy_in map of element 0 with nods = [3,4,5,7,9,8,1] (user order):
    #   3 -> 0; 4 -> fs+0 = 10; 5 -> 11; 7 -> 1,2,3; 9 -> 12,13,14;
    #   8 -> 4,5,6; 1 -> 7,8,9.
    d = {
        "elements": [[3, 4, 5, 7, 9, 8, 1], [4, 6, 5, 10, 11, 9, 2]],
        "n_dof": {"3": 1, "4": 1, "5": 1, "6": 1, "7": 3, "8": 3, "9": 3, "10": 3, "11": 3, "1": 3, "2": 3},
        "offset": {"3": 0, "7": 1, "8": 4, "1": 7, "6": 10, "10": 11, "11": 14, "2": 17, "4": 20, "5": 21, "9": 22},
        "fs": [10, 10, 5], "ss": [5, 5, 0], "generation": [0, 0, -1],
        "factor": [[[3, 0], [7, 1], [8, 4], [1, 7]], [[6, 0], [10, 1], [11, 4], [2, 7]], [[4, 0], [5, 1], [9, 2]]],
        "schur": [[[4, 0], [5, 1], [9, 2]], [[4, 0], [5, 1], [9, 2]], []],
        "mapper": [[[0, 0, 0, 5]], [[0, 0, 0, 5]], []],
        "copy_in_map_e0": [0, 10, 11, 1, 2, 3, 12, 13, 14, 4, 5, 6, 7, 8, 9],
        "ndof": 25,
    }
    with open(os.path.join(HERE, "ex1_maps.json"), "w") as f:
        json.dump(d, f, indent=1)


if __name__ == "__main__":
    make_fronts()
    make_q2()
    make_ex1()
    print("fixtures written to", HERE)
