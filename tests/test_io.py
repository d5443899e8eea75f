"""On-disk formats: MatrixMarket round trip, pattern fingerprint, plan cache (CPU only)."""
import ctypes as C

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import ispm_amd as I
import ispm_amd.gen as G


def test_mm_roundtrip(tmp_path):
    A = G.random_sparse(300, 2, max_len=9)
    p = tmp_path / "a.mtx"
    I.mm_write(p, A)
    B = I.mm_read(p)
    assert abs(B.to_scipy() - A.to_scipy()).max() == 0
    # scipy reads what we write
    S = scipy.io.mmread(str(p)).tocsr()
    assert abs(S - A.to_scipy()).max() == 0


def test_mm_symmetric_and_pattern(tmp_path):
    M = sp.random(50, 50, density=0.1, random_state=3, format="csr")
    M = (M + M.T + sp.eye(50)).tocsr()
    p = tmp_path / "s.mtx"
    scipy.io.mmwrite(str(p), M, symmetry="symmetric")
    B = I.mm_read(p)
    assert abs(B.to_scipy() - M).max() < 1e-15
    q = tmp_path / "p.mtx"
    q.write_text("%%MatrixMarket matrix coordinate pattern general\n3 4 3\n1 1\n2 4\n3 2\n")
    P = I.mm_read(q)
    assert P.n_cols == 4 and P.nnz == 3 and np.all(P.vals == 1.0)


def test_mm_rejects_bad_files(tmp_path):
    q = tmp_path / "bad.mtx"
    q.write_text("not a matrix\n")
    with pytest.raises(I.IspmError):
        I.mm_read(q)
    with pytest.raises(I.IspmError):
        I.mm_read(tmp_path / "missing.mtx")


def test_fingerprint_depends_on_pattern_only():
    A = G.poisson7(8)
    B = I.CSR(A.n_rows, A.n_cols, A.row_ptr, A.cols, A.vals * 3)
    assert I.fingerprint(A) == I.fingerprint(B)
    cols = A.cols.copy()
    cols[5], cols[6] = cols[6], cols[5]
    C2 = I.CSR(A.n_rows, A.n_cols, A.row_ptr, cols, A.vals)
    assert I.fingerprint(A) != I.fingerprint(C2)


def test_plan_cache_store_lookup(tmp_path):
    path = str(tmp_path / "plans.txt").encode()
    L = I.lib()
    lp = I.Launch(256, 64, 2, 0, 0)
    I.check(L.ispm_plan_cache_store(path, 0xABCDEF, I.F64, I.F32, 64, 1, 0, 1, C.byref(lp), 1.5e-4))
    lp2 = I.Launch(512, 128, 1, 0, 0)
    I.check(L.ispm_plan_cache_store(path, 0xABCDEF, I.F64, I.F32, 128, 2, 1, 0, C.byref(lp2), 1.0e-4))
    S, W, V, D = C.c_uint(), C.c_uint(), C.c_int(), C.c_int()
    out, sec = I.Launch(), C.c_double()
    I.check(L.ispm_plan_cache_lookup(path, 0xABCDEF, I.F64, I.F32, C.byref(S), C.byref(W), C.byref(V),
                                     C.byref(D), C.byref(out), C.byref(sec)))
    assert (S.value, W.value, V.value, D.value, out.N, out.S, out.H) == (128, 2, 1, 0, 512, 128, 1)
    assert L.ispm_plan_cache_lookup(path, 0x1234, I.F64, I.F32, C.byref(S), C.byref(W), C.byref(V),
                                    C.byref(D), C.byref(out), C.byref(sec)) != 0


@pytest.mark.parametrize("size_line, code", [
    ("3 3 -1", "INVALID"),                 # negative entry count
    ("3 3 3000000000", "UNSUPPORTED"),     # beyond 32-bit row pointers
])
def test_mm_rejects_bad_entry_counts(tmp_path, size_line, code):
    q = tmp_path / "count.mtx"
    q.write_text(f"%%MatrixMarket matrix coordinate real general\n{size_line}\n1 1 1.0\n")
    with pytest.raises(I.IspmError) as e:
        I.mm_read(q)
    assert e.value.code == getattr(I, "ERR_" + code)


def test_mm_symmetric_count_doubles_before_the_limit(tmp_path):
    # 1.5e9 stored entries fit 32 bits, their mirrors do not: refused before reading any
    q = tmp_path / "sym.mtx"
    q.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 1500000000\n1 1 1.0\n")
    with pytest.raises(I.IspmError) as e:
        I.mm_read(q)
    assert e.value.code == I.ERR_UNSUPPORTED


def test_mm_truncated_file_is_an_error(tmp_path):
    q = tmp_path / "short.mtx"
    q.write_text("%%MatrixMarket matrix coordinate real general\n3 3 4\n1 1 1.0\n2 2 1.0\n")
    with pytest.raises(I.IspmError) as e:
        I.mm_read(q)
    assert e.value.code == I.ERR_INVALID
