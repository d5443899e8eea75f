"""Product fastainv_sym (libispm0) against the restated oracle: bit-identical factors."""
import numpy as np
import pytest

import ispm_amd as I
import ispm_amd.gen as G
import oracle as O


@pytest.mark.parametrize("out_type", [I.F32, I.F64])
@pytest.mark.parametrize("droptol", [0.0, 0.01, 0.05, 0.125])
@pytest.mark.parametrize("name", ["poisson", "random", "contrast"])
def test_product_matches_oracle(name, droptol, out_type):
    A = {"poisson": lambda: G.poisson7(8),
         "random": lambda: G.random_sparse(400, 31, max_len=9, spd=True),
         "contrast": lambda: G.stencil27_contrast(8, block=4)}[name]()
    p1, p2 = I.fastainv_sym(A, out_type, droptol=droptol)
    o = O.fastainv_sym(A, out_double=out_type == I.F64, droptol=droptol)
    for prod, orc in ((p1, o[0]), (p2, o[1])):
        np.testing.assert_array_equal(prod.row_ptr, orc[0])
        np.testing.assert_array_equal(prod.cols, orc[1])
        np.testing.assert_array_equal(prod.vals.astype(np.float64), orc[2])


@pytest.mark.parametrize("max_elts", [8, 64, 400, 4000])
def test_truncation_matches_oracle(max_elts):
    A = G.poisson7(7)
    p1, p2 = I.fastainv_sym(A, I.F64, max_elts=max_elts, droptol=0.0)
    o = O.fastainv_sym(A, out_double=True, max_elts=max_elts, droptol=0.0)
    np.testing.assert_array_equal(p1.cols, o[0][1])
    np.testing.assert_array_equal(p1.vals, o[0][2])
    np.testing.assert_array_equal(p2.vals, o[1][2])


def test_breakdown_zero_pivot():
    # leading entry 0: D_00 = 0 -> breakdown, D[0] = 0 and column 0 updates nothing
    A = I.CSR(3, 3, [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2], [0.0, 1.0, 1.0, 4.0, 1.0, 1.0, 3.0])
    p1, p2 = I.fastainv_sym(A, I.F64, droptol=0.0)
    o = O.fastainv_sym(A, out_double=True, droptol=0.0)
    np.testing.assert_array_equal(p1.vals, o[0][2])
    # Ainv1 row n-1 is column 0 of W scaled by sqrt|D_0| = 0
    assert p1.vals[p1.row_ptr[2]:p1.row_ptr[3]].tolist() == [0.0]


def test_invalid_n():
    with pytest.raises(I.IspmError):
        I.fastainv_sym(I.CSR(0, 0, [0], np.zeros(0, np.int32), np.zeros(0)), I.F64)


def test_progress_callback_called_per_column():
    A = G.poisson7(5)
    calls = []
    I.fastainv_sym(A, I.F32, progress=lambda v: calls.append(1))
    assert len(calls) == A.n_rows
