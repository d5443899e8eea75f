"""Host sparse product (ispm_csr_multiply) against scipy, and the product preconditioner
choice: C = A * B in FP64, sorted columns, same pattern as scipy's product including entries
that cancel to zero."""
import numpy as np
import pytest
import scipy.sparse as sp

import ispm_amd as I
import ispm_amd.gen as G
from ispm_amd.solver import preconditioner_matrices


def dense(A):
    return A.to_scipy().toarray()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_multiply_matches_scipy(seed):
    A = G.random_sparse(300, seed, min_len=0, max_len=12, n_cols=250)
    B = G.random_sparse(250, seed + 10, min_len=0, max_len=9, n_cols=400)
    C = I.csr_multiply(A, B)
    ref = (A.to_scipy() @ B.to_scipy()).tocsr()
    ref.sort_indices()
    assert (C.n_rows, C.n_cols) == ref.shape
    np.testing.assert_array_equal(C.row_ptr, ref.indptr)
    np.testing.assert_array_equal(C.cols, ref.indices)
    np.testing.assert_allclose(C.vals, ref.data, rtol=1e-14, atol=1e-300)


def test_multiply_integer_exact_and_threads():
    """Above the threading threshold, integer values: exact values; the pattern is the
    symbolic product (entries that cancel to zero stay, unlike scipy's product)."""
    A = G.random_sparse(20000, 5, min_len=1, max_len=6, integer=True)
    B = G.random_sparse(20000, 6, min_len=1, max_len=6, integer=True)
    C = I.csr_multiply(A, B)
    ones = lambda M: sp.csr_matrix((np.ones(M.nnz), M.cols, M.row_ptr), shape=(M.n_rows, M.n_cols))
    sym = (ones(A) @ ones(B)).tocsr()
    sym.sort_indices()
    np.testing.assert_array_equal(C.row_ptr, sym.indptr)
    np.testing.assert_array_equal(C.cols, sym.indices)
    ref = (A.to_scipy() @ B.to_scipy()).tocsr()
    got = C.to_scipy()
    assert abs(got - ref).max() == 0
    assert (C.vals == 0).sum() == sym.nnz - ref.count_nonzero()


def test_multiply_fp32_output_and_errors():
    A = G.random_sparse(50, 7, min_len=0, max_len=5)
    C = I.csr_multiply(A, A, I.F32)
    assert C.vals.dtype == np.float32
    np.testing.assert_allclose(dense(C), dense(A) @ dense(A), rtol=1e-6, atol=1e-6)
    B = G.random_sparse(40, 8, min_len=0, max_len=5)
    with pytest.raises(I.IspmError):
        I.csr_multiply(A, B)          # 50 columns times 40 rows


def test_product_preconditioner_choice():
    A = G.poisson7(12)
    a1, a2 = I.fastainv_sym(A, I.F64, droptol=0.125)
    mats = preconditioner_matrices((a1, a2), True)
    assert len(mats) == 1
    np.testing.assert_allclose(dense(mats[0]), dense(a2) @ dense(a1), rtol=1e-13, atol=1e-14)
    assert len(preconditioner_matrices((a1, a2), False)) == 2
    auto = preconditioner_matrices((a1, a2))
    P = I.csr_multiply(a2, a1)
    assert len(auto) == (1 if P.nnz < 0.9 * (a1.nnz + a2.nnz) else 2)
