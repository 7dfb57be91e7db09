"""Whole small fronts in one launch (uhmk_chol_front_small, tuning key 9):
on Cholesky tree levels whose fronts all have fs <= 64 one workgroup per
front runs the assembly, diagonal block, panel TRSM and Schur complement
computations of the four level launches in the same order, so every front
must come out bit for bit as the four launches leave it, for FP64 fronts
(FP64 products) and FP32 fronts (FP32 products), and the solution must
still reach the oracle-pinned accuracy.  The meshes cover Q3 hexahedra
(leaf fronts of 64 entries, fs 8..27), Q4 hexahedra (n = 125, fs up to
64, Schur blocks wider than one 64-tile) and 2D variable-order quads
(ragged n, fronts with fs = 0 on the boundary)."""
import numpy as np
import pytest

import uhm_amd as U

pytestmark = pytest.mark.gpu

MESHES = [
    ("hex-q3-8cube", 2, (8, 8, 8), 3, 0),
    ("hex-q4-5cube", 2, (5, 5, 5), 4, 0),
    ("quad-varp-48", 3, (48, 48), 6, 777),
]


def _mesh(kind, dims, p, seed):
    m = U.Mesh()
    m.build_synthetic(kind, dims, p, seed)
    m.lock()
    m.create_matrix_without_buffer(U.UHM_REAL, 1)
    m.create_matrix_buffer()
    m.set_random_params(seed=11, alpha=-1.0, beta=0.0, scale_kind=1, p0=0.5)
    m.random_matrix()
    m.set_rhs()
    return m


@pytest.mark.parametrize("prec", [U.PREC_FP64, U.PREC_MIXED])
@pytest.mark.parametrize("case", MESHES, ids=[c[0] for c in MESHES])
def test_fused_small_fronts_match_level_launches(lib, case, prec):
    name, kind, dims, p, seed = case
    print("case", name, "precision mode", prec, flush=True)
    m = _mesh(kind, dims, p, seed)
    if prec:
        m.set_precision(prec, 0)
    m.set_profile(True)
    m.set_tuning(9, 0)
    assert m.chol_with_free() == 0
    assert m.kernel_stats()["chol_small"][3] == 0
    ref = [m.fetch_front(f) for f in range(m.n_fronts())]
    m.set_tuning(9, 1)
    assert m.chol_with_free() == 0
    stats = m.kernel_stats()
    fused_levels = stats["chol_small"][3]
    small = [f for f in range(m.n_fronts()) if m.front_info(f)["fs"] <= 64]
    print("  fused launches %d, fronts with fs <= 64: %d of %d" % (fused_levels, len(small), m.n_fronts()), flush=True)
    assert fused_levels >= 1
    # the factor as stored: the lower triangle (L, L21, the Schur complement)
    # and the fs x fs diagonal block (its strict upper part holds the
    # inverse), compared as bit patterns; the rest of the upper triangle is
    # not part of a Cholesky front
    def stored(F, fs):
        keep = np.tril(np.ones(F.shape, bool))
        for b0 in range(0, fs, 64):
            b1 = min(b0 + 64, fs)
            keep[b0:b1, b0:b1] = True
        return F.view(np.uint64)[keep]

    differ = []
    for f in range(m.n_fronts()):
        fs = m.front_info(f)["fs"]
        if not np.array_equal(stored(m.fetch_front(f), fs), stored(ref[f], fs)):
            differ.append(f)
    assert not differ, differ[:10]
    m.solve_chol()
    n = m.improve(U.UHM_CHOL, 1e-12, 20) if prec else 0
    eta = m.backward_error()
    print("  corrections %d, eta %.2e" % (n, eta), flush=True)
    assert eta <= (1e-12 if prec else 1e-13)


def test_fused_small_threshold(lib):
    """Off by default; key 9 = 512 fuses exactly the levels of at least 512
    fronts whose fronts all have fs <= 64."""
    m = _mesh(2, (8, 8, 8), 3, 0)
    assert m.get_tuning(9) == 0
    m.set_tuning(9, 512)
    m.set_profile(True)
    assert m.chol_with_free() == 0
    levels = {}
    for f in range(m.n_fronts()):
        fi = m.front_info(f)
        levels.setdefault(fi["level"], []).append(fi["fs"])
    want = sum(1 for v in levels.values() if len(v) >= 512 and max(v) <= 64)
    assert want >= 1
    assert m.kernel_stats()["chol_small"][3] == want
    with pytest.raises(U.UHMError):
        m.set_tuning(9, -1)
