"""lduMatrix input (ispm_ldu_to_csr, ispm_ldu_values) and the DroptolTuner C-ABI, on the CPU:
the conversion against the oracle's entry-by-entry dense matrix, the coefficient positions,
argument checks, and the tuner against the oracle's tuner on the same time sequences."""
import numpy as np
import pytest

import ispm_amd as I
import ispm_amd.gen as G
from ispm_amd import foam as F
import oracle as O


def _ldu(nx, ny, nz, asym=0.0, seed=4001):
    return F.LduMatrix(*G.ldu_mesh(nx, ny, nz, seed=seed, asymmetry=asym))


@pytest.mark.parametrize("asym", [0.0, 0.5])
@pytest.mark.parametrize("shape", [(4, 3, 5), (1, 1, 7), (9, 1, 1), (1, 1, 1)])
def test_ldu_to_csr_matches_dense(shape, asym):
    L = _ldu(*shape, asym=asym)
    A, pos = F.ldu_to_csr(L)
    D = O.ldu_dense(L.n, L.lower_addr, L.upper_addr, L.diag, L.upper, L.lower)
    np.testing.assert_array_equal(A.to_scipy().toarray(), D)
    assert A.nnz == L.n + 2 * L.n_faces
    rp, cols = np.asarray(A.row_ptr), np.asarray(A.cols)
    for i in range(L.n):
        assert np.all(np.diff(cols[rp[i]:rp[i + 1]]) > 0)      # columns ascending
    # positions: every coefficient lands where pos says
    np.testing.assert_array_equal(np.sort(pos), np.arange(A.nnz))
    src = np.concatenate([L.diag, L.upper, L.upper if L.lower is None else L.lower])
    np.testing.assert_array_equal(np.asarray(A.vals)[pos], src)


def test_ldu_faces_in_any_order():
    L = _ldu(5, 4, 3, asym=0.3)
    perm = np.random.default_rng(0).permutation(L.n_faces)
    Lp = F.LduMatrix(L.n, L.lower_addr[perm], L.upper_addr[perm], L.diag, L.upper[perm], L.lower[perm])
    A, _ = F.ldu_to_csr(L)
    B, _ = F.ldu_to_csr(Lp)
    np.testing.assert_array_equal(A.to_scipy().toarray(), B.to_scipy().toarray())


def test_ldu_values_refresh():
    L = _ldu(6, 5, 4, asym=0.2)
    A, pos = F.ldu_to_csr(L)
    L2 = _ldu(6, 5, 4, asym=0.2, seed=77)           # same addressing, new coefficients
    v = F.ldu_values(L2, pos)
    A2, _ = F.ldu_to_csr(L2)
    np.testing.assert_array_equal(v, np.asarray(A2.vals))


@pytest.mark.parametrize("bad", ["order", "range", "negative", "duplicate"])
def test_ldu_to_csr_rejects(bad):
    la, ua = np.array([0, 1], np.int32), np.array([1, 2], np.int32)
    if bad == "order":
        la, ua = np.array([1, 1], np.int32), np.array([0, 2], np.int32)
    elif bad == "range":
        ua = np.array([1, 3], np.int32)
    elif bad == "negative":
        la = np.array([-1, 1], np.int32)
    else:
        la, ua = np.array([0, 0], np.int32), np.array([1, 1], np.int32)
    with pytest.raises(I.IspmError):
        F.ldu_to_csr(F.LduMatrix(3, la, ua, np.ones(3), -np.ones(2)))


def test_ldu_empty_mesh():
    A, pos = F.ldu_to_csr(F.LduMatrix(2, np.zeros(0, np.int32), np.zeros(0, np.int32), np.array([2.0, 3.0]),
                                      np.zeros(0)))
    np.testing.assert_array_equal(A.to_scipy().toarray(), np.diag([2.0, 3.0]))
    np.testing.assert_array_equal(pos, [0, 1])


@pytest.mark.parametrize("seed", range(4))
def test_droptol_tuner_cabi_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    a, b = F.DroptolTuner(), O.DroptolTuner()
    for step in range(30):
        assert a.droptol() == b.get()
        opt = rng.uniform(-9, -4)
        t = (np.log2(a.droptol()) - opt) ** 2 + rng.uniform(0, 0.1)
        a.note_time(t)
        b.note(t)
        assert a.best() == b.best() and a.iterations() == b.iterations()
        if step == 15:
            a.restart()
            b.restart()


def test_droptol_tuner_rejects_empty_range():
    with pytest.raises(I.IspmError):
        F.DroptolTuner(-3.0, -10.0)
