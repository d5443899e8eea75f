"""Multi-RHS block PCG (SpMM-based) and the persistent batched solver for many small SPD
systems, checked with the CPU oracle's FP64 residuals."""
import numpy as np
import pytest

import ispm_amd as I
import ispm_amd.gen as G
import oracle as O
from ispm_amd import batch as B

pytestmark = pytest.mark.gpu


def relres(A, b, x):
    return float(np.linalg.norm(b - O.csr_spmv(A, x)) / np.linalg.norm(b))


@pytest.mark.parametrize("k", [1, 3, 8])
def test_block_pcg_matches_single(gpu, k):
    A = G.fv7_lognormal(16, rcm=True)
    M = I.fastainv_sym(A, I.F64, droptol=0.05)
    Bm = np.stack([O.csr_spmv(A, G.uniform(100 + j, A.n_rows)) for j in range(k)])
    X, stats = B.block_solve(A, Bm, M=M, tol=1e-10)
    for j in range(k):
        assert stats[j]["converged"]
        assert relres(A, Bm[j], X[j]) <= 1e-10
    # each column follows the same recurrence as a single-RHS PCG on that column
    from ispm_amd import solver as S
    r0 = S.solve(A, Bm[0], M=M, tol=1e-10, precision="fp64", method=S.PCG)
    assert abs(stats[0]["iterations"] - r0.iterations) <= 1


@pytest.mark.parametrize("k", [2, 5])
def test_block_pcg_symmetric_storage(gpu, k):
    """Block PCG with A as a symmetric diagonal plan (its SpMM reads the upper planes and
    the mirrors once for all columns): every column converges to the oracle's residual, in
    the iterations of the same block solve with A stored in full (within one)."""
    A = G.stencil27_contrast(16, block=4)
    M = I.fastainv_sym(A, I.F64, droptol=0.05)
    Bm = np.stack([O.csr_spmv(A, G.uniform(700 + j, A.n_rows)) for j in range(k)])
    Xs, ss = B.block_solve(A, Bm, M=M, tol=1e-10, symmetric=True)
    Xf, sf = B.block_solve(A, Bm, M=M, tol=1e-10)
    for j in range(k):
        assert ss[j]["converged"] and relres(A, Bm[j], Xs[j]) <= 1e-10
        assert abs(ss[j]["iterations"] - sf[j]["iterations"]) <= 1


def test_block_graph_and_eager_agree(gpu):
    """Batches replayed from the captured hipGraph (several replays, with the host's checks
    and a restart in between) give the eager launches' X bit for bit and the same counts."""
    A = G.fv7_lognormal(16, rcm=True)
    M = I.fastainv_sym(A, I.F64, droptol=0.05)
    Bm = np.stack([O.csr_spmv(A, G.uniform(800 + j, A.n_rows)) for j in range(4)])
    Xg, sg = B.block_solve(A, Bm, M=M, tol=1e-10, check_every=8, use_graph=True)
    Xe, se = B.block_solve(A, Bm, M=M, tol=1e-10, check_every=8, use_graph=False)
    np.testing.assert_array_equal(Xg, Xe)
    for j in range(4):
        assert sg[j]["iterations"] == se[j]["iterations"] and sg[j]["replacements"] == se[j]["replacements"]
        assert sg[j]["converged"] and relres(A, Bm[j], Xg[j]) <= 1e-10


def test_block_maxit_stop_applies_pending_updates(gpu):
    """Stopped by maxit, every column's X holds all its updates (the last one is applied by
    the flush, not by a next iteration): it matches a single-RHS PCG stopped after as many
    updates, within rounding."""
    from ispm_amd import solver as S
    A = G.fv7_lognormal(16, rcm=True)
    M = I.fastainv_sym(A, I.F64, droptol=0.05)
    Bm = np.stack([O.csr_spmv(A, G.uniform(950 + j, A.n_rows)) for j in range(3)])
    X, st = B.block_solve(A, Bm, M=M, tol=1e-14, maxit=8, check_every=4)
    assert all(s_["iterations"] == 8 and not s_["converged"] for s_ in st)
    for j in range(3):
        ref = None
        for m in (7, 8, 9):
            r = S.solve(A, Bm[j], M=M, tol=1e-14, precision="fp64", method=S.PCG, maxit=m, check_every=4,
                        fsai_product=False, split_diag=False, symmetric=False)
            if r.iterations == 8:
                ref = r
                break
        assert ref is not None, [S.solve(A, Bm[j], M=M, tol=1e-14, precision="fp64", method=S.PCG, maxit=m,
                                         check_every=4, fsai_product=False, split_diag=False,
                                         symmetric=False).iterations for m in (7, 8, 9)]
        scale = np.max(np.abs(ref.x))
        assert np.max(np.abs(X[j] - ref.x)) <= 1e-9 * scale


def test_block_columns_converge_independently(gpu):
    A = G.poisson7(12)
    Bm = np.stack([O.csr_spmv(A, G.uniform(5, A.n_rows)), np.zeros(A.n_rows) + 1e-3,
                   O.csr_spmv(A, np.ones(A.n_rows))])
    X, stats = B.block_solve(A, Bm, M=None, tol=1e-9)
    for j in range(3):
        assert stats[j]["converged"] and relres(A, Bm[j], X[j]) <= 1e-9


def small_systems(count, seed, lo=200, hi=3000):
    sizes = np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * G.uniform01(seed, count)).astype(int)
    ks = 10 + (G.uniform01(seed + 1, count) * 17).astype(int)
    return [G.rgg_laplacian(int(n), int(k), seed=seed + 10 + i, dtype=np.float32) for i, (n, k) in enumerate(zip(sizes, ks))]


def test_batched_persistent_solver(gpu):
    systems = small_systems(40, 7)
    rhs = [O.csr_spmv(A, G.uniform(900 + i, A.n_rows)).astype(np.float32) for i, A in enumerate(systems)]
    xs, iters, rel = B.batch_solve(systems, rhs, tol=1e-6, maxit=5000)
    assert np.all(iters > 0) and np.all(iters < 5000)
    for A, b, x in zip(systems, rhs, xs):
        A64 = I.CSR(A.n_rows, A.n_cols, A.row_ptr, A.cols, A.vals.astype(np.float64))
        assert relres(A64, b.astype(np.float64), x.astype(np.float64)) <= 1e-5


def test_batched_matches_single_system_runs(gpu):
    systems = small_systems(6, 11)
    rhs = [O.csr_spmv(A, G.uniform(50 + i, A.n_rows)).astype(np.float32) for i, A in enumerate(systems)]
    xs_all, it_all, _ = B.batch_solve(systems, rhs, tol=1e-6)
    for i in range(len(systems)):
        xs1, it1, _ = B.batch_solve([systems[i]], [rhs[i]], tol=1e-6)
        assert it1[0] == it_all[i]
        np.testing.assert_array_equal(xs1[0], xs_all[i])
