"""The workload configurations at their stated sizes, checked with the CPU oracle.

cfg1  7-point Poisson 32^3 (diagonal 6+1e-3), default plan, FSAI-PCG to 1e-8.
cfg2  128^3 finite-volume 7-point operator, lognormal(0, 1.5) face conductances, randomly
      permuted then RCM-reordered: FP64 SpMV per-entry bound and pipelined FSAI-PCG to 1e-8.
cfg4a the cfg2 matrix with 3 and 8 right-hand sides solved as one block PCG.
cfg4b 512 random-geometric-graph Laplacians (1k..64k rows, log-uniform, k = 10..26
      neighbours, shift 1e-2, FP32) in one persistent launch.
(cfg3, 27-point 192^3, is covered in test_gpu_spmv.py and by the benchmark.)"""
import numpy as np
import pytest

import ispm_amd as I
import ispm_amd.gen as G
import oracle as O
from ispm_amd import batch as B
from ispm_amd import solver as S

pytestmark = pytest.mark.gpu

_CACHE = {}


def cfg2():
    if "A" not in _CACHE:
        _CACHE["A"] = G.fv7_lognormal(128)
    return _CACHE["A"]


def relres(A, b, x):
    return float(np.linalg.norm(b - O.csr_spmv(A, x)) / np.linalg.norm(b))


def test_cfg1_poisson_32(gpu):
    A = G.poisson7(32)
    assert A.n_rows == 32768 and A.nnz == 223232
    p = I.Plan(A, I.F64)
    x = G.uniform(0, A.n_rows)
    n = p.n_rows_padded
    xd = I.DeviceArray.from_numpy(np.concatenate([x, np.zeros(n - A.n_rows)]))
    yd = I.DeviceArray(n)
    p.execute(xd, yd)
    y = yd.numpy()[:A.n_rows]
    assert np.all(np.abs(y - O.csr_spmv(A, x)) <= 1e-13 * O.csr_spmv_abs(A, x))
    b = O.csr_spmv(A, G.uniform(1, A.n_rows))
    M = I.fastainv_sym(A, I.F64)
    for prec, meth in (("fp64", S.PPCG), ("fp64", S.PCG), ("mixed", S.PCG)):
        res = S.solve(A, b, M=M, tol=1e-8, precision=prec, method=meth)
        assert res.converged and relres(A, b, res.x) <= 1e-8, (prec, meth)


def test_cfg2_spmv_and_pipelined_fsai_pcg(gpu):
    A = cfg2()
    assert A.n_rows == 2097152 and A.nnz == 14581760
    p = I.Plan(A, I.F64)
    x = G.uniform(0, A.n_rows)
    n = p.n_rows_padded
    xd = I.DeviceArray.from_numpy(np.concatenate([x, np.zeros(n - A.n_rows)]))
    yd = I.DeviceArray(n)
    p.execute(xd, yd)
    y = yd.numpy()[:A.n_rows]
    assert np.all(np.abs(y - O.csr_spmv(A, x)) <= 1e-13 * O.csr_spmv_abs(A, x))
    b = O.csr_spmv(A, G.uniform(1, A.n_rows))
    M = I.fastainv_sym(A, I.F64, max_elts=64 * A.n_rows)
    res = S.solve(A, b, M=M, tol=1e-8, precision="fp64", method=S.PPCG)
    assert res.converged and relres(A, b, res.x) <= 1e-8


@pytest.mark.parametrize("k", [3, 8])
def test_cfg4a_block_rhs(gpu, k):
    A = cfg2()
    M = I.fastainv_sym(A, I.F64, max_elts=64 * A.n_rows)
    Bm = np.stack([O.csr_spmv(A, G.uniform(300 + j, A.n_rows)) for j in range(k)])
    X, stats = B.block_solve(A, Bm, M=M, tol=1e-8)
    for j in range(k):
        assert stats[j]["converged"] and relres(A, Bm[j], X[j]) <= 1e-8


def test_cfg4b_512_systems(gpu):
    count = 512
    sizes = np.exp(np.log(1000) + (np.log(64000) - np.log(1000)) * G.uniform01(4001, count)).astype(int)
    ks = 10 + np.minimum(16, (G.uniform01(4002, count) * 17).astype(int))
    systems = [G.rgg_laplacian(int(n), int(k), seed=5000 + i, dtype=np.float32)
               for i, (n, k) in enumerate(zip(sizes, ks))]
    rhs = [O.csr_spmv(A, G.uniform(6000 + i, A.n_rows)).astype(np.float32) for i, A in enumerate(systems)]
    xs, iters, rel = B.batch_solve(systems, rhs, tol=1e-6, maxit=20000)
    assert np.all(iters > 0) and np.all(iters < 20000)
    worst = 0.0
    for A, b, x in zip(systems, rhs, xs):
        A64 = I.CSR(A.n_rows, A.n_cols, A.row_ptr, A.cols, A.vals.astype(np.float64))
        worst = max(worst, relres(A64, b.astype(np.float64), x.astype(np.float64)))
    assert worst <= 1e-5
