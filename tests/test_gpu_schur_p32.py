"""The Cholesky updates of FP64 fronts with FP32 MFMA on an FP32 copy of
their operand (tuning key 15 = 1, the default: uhmk_p32_pack +
uhmk_gemm_update_p32 for the trailing updates after each panel group and
the final syrk of chol_hier, LINAL/src/dense/chol.cxx:86-90,
UHM/src/matrix/uhm/fla/chol/decompose.cxx:95-96) against the same products
on element loads of the front (key 15 = 0).

Both round every operand entry once to FP32 and run the same MFMA sequence
per tile, so every front's factor data is the same bits: its lower triangle
and W's strict lower part, which sits transposed in the strict upper
triangle of each 64 x 64 diagonal block (the rest of the upper triangle is
never read on the Cholesky path and keeps what earlier factorizations left
there); the copy path must actually run
(its launches appear in the kernel statistics) and the mixed-precision
refinement still reaches 1e-12."""
import sys

import numpy as np
import pytest

import uhm_amd as U

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def test_schur_p32_same_bits(lib):
    # FP32 products on the FP64 top levels (precision mode 1; in modes 2 / 3
    # those levels take the packed BF16 / FP16 products when their Schur
    # blocks are large, as here)
    prec = U.PREC_MIXED
    m = U.Mesh()
    m.build_synthetic(2, (10, 10, 9), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=13, alpha=64.0, beta=0.0, scale_kind=1, p0=0.5)
    m.random_matrix()
    m.set_rhs()
    m.set_precision(prec, 2)
    assert m.get_tuning(15) == 1
    m.set_tuning(15, 0)
    assert m.chol_with_free() == 0
    F0 = [m.fetch_front(f) for f in range(m.n_fronts())]
    m.set_tuning(15, 1)
    m.set_profile(True)
    assert m.chol_with_free() == 0
    st = m.kernel_stats()
    m.set_profile(False)
    say("  p32 launches: pack %d, Schur products %d, trailing updates %d" % (st["p32_pack"][3], st["schur_p32"][3],
                                                                          st["gemm_p32"][3]))
    assert st["schur_p32"][3] > 0 and st["gemm_p32"][3] > 0
    assert st["p32_pack"][3] == st["schur_p32"][3] + st["gemm_p32"][3]
    for f in range(m.n_fronts()):
        a, b = F0[f], m.fetch_front(f)
        fs = m.front_info(f)["fs"]
        used = np.tril(np.ones(a.shape, bool))
        for c0 in range(0, fs, 64):
            c1 = min(c0 + 64, fs)
            used[c0:c1, c0:c1] = True
        assert np.array_equal(a.view(np.uint64)[used], b.view(np.uint64)[used]), f
    m.solve_chol()
    m.improve(U.UHM_CHOL, 1e-12, 10)
    say("  refinement:", m.refine_history())
    assert m.backward_error() <= 1e-12


def test_p32_graph_replays_match_eager(lib):
    """The copy launches inside a captured factorization: every replay (with
    eager work between them) gives the eager factor data and solution."""
    m = U.Mesh()
    m.build_synthetic(2, (10, 10, 9), 3, 0)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=17, alpha=64.0, beta=0.0, scale_kind=1, p0=0.5)
    m.random_matrix()
    m.set_rhs()
    m.set_precision(U.PREC_MIXED, 2)

    def run():
        assert m.chol_with_free() == 0
        m.solve_chol()
        out = []
        for f in range(m.n_fronts()):
            a = m.fetch_front(f)
            fs = m.front_info(f)["fs"]
            used = np.tril(np.ones(a.shape, bool))
            for c0 in range(0, fs, 64):
                c1 = min(c0 + 64, fs)
                used[c0:c1, c0:c1] = True
            out.append(a.view(np.uint64)[used].copy())
        return out, m.fetch_x().copy()

    F0, x0 = run()
    m.set_graph(True)
    for rep in range(3):
        F1, x1 = run()
        for f in range(m.n_fronts()):
            assert np.array_equal(F0[f], F1[f]), (rep, f)
        assert np.array_equal(x0.view(np.uint64), x1.view(np.uint64)), rep
        m.check_chol()
        say("  replay %d: backward error %.2e" % (rep, m.backward_error()))
    m.set_graph(False)
