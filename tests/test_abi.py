"""The C-ABI library loads and exports every entry point declared in include/ispm.h (no GPU
work is done here)."""
import ctypes
import os
import re

import numpy as np

import ispm_amd as I

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "ispm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ispm_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    names = declared()
    assert len(names) > 40
    lib = ctypes.CDLL(I.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared()) <= set(I.declared_symbols())


def test_static_archive_built():
    assert os.path.exists(os.path.join(I.PKG_DIR, "libispm0-pic.a"))


def test_cpu_only_calls():
    # host-side conversion works without a device
    A = I.CSR(2, 2, [0, 1, 2], [0, 1], [1.0, 2.0])
    hs = I.csr_to_slell(A, 32)
    assert hs.n_rows == 512 and hs.n_slices == 16


def test_split_diagonal_plan_argument_checks():
    """Split-diagonal plans need a square matrix and the device; both are checked before any
    device work."""
    import pytest
    A = I.CSR(2, 3, [0, 1, 2], [0, 2], [1.0, 2.0])
    with pytest.raises(I.IspmError):
        I.Plan(A, I.F64, flags=I.SPMV_PLAN_SPLIT_DIAG)
    B = I.CSR(2, 2, [0, 1, 2], [0, 1], [1.0, 2.0])
    with pytest.raises(I.IspmError):
        I.Plan(B, I.F64, flags=I.SPMV_PLAN_SPLIT_DIAG | I.SPMV_PLAN_NO_DEVICE)


def test_symmetric_plan_argument_checks():
    """Symmetric diagonal plans need an exactly symmetric square matrix on at most 16 upper
    diagonals, and the device; the checks run before any device work."""
    import pytest
    B = I.CSR(3, 3, [0, 2, 3, 4], [0, 1, 1, 2], [2.0, 1.0, 2.0, 2.0])   # (0,1) without (1,0)
    with pytest.raises(I.IspmError):
        I.Plan(B, I.F64, flags=I.SPMV_PLAN_SYMMETRIC)
    C_ = I.CSR(2, 2, [0, 2, 4], [0, 1, 0, 1], [2.0, 1.0, 1.5, 2.0])   # (0,1) != (1,0)
    with pytest.raises(I.IspmError):
        I.Plan(C_, I.F64, flags=I.SPMV_PLAN_SYMMETRIC)
    n = 40
    D = I.CSR(n, n, np.arange(n + 1) * n, np.tile(np.arange(n, dtype=np.int32), n), np.ones(n * n))  # dense: 40 diagonals
    with pytest.raises(I.IspmError):
        I.Plan(D, I.F64, flags=I.SPMV_PLAN_SYMMETRIC)
    E_ = I.CSR(2, 2, [0, 2, 4], [0, 1, 0, 1], [2.0, 1.0, 1.0, 2.0])   # symmetric, but no device
    with pytest.raises(I.IspmError):
        I.Plan(E_, I.F64, flags=I.SPMV_PLAN_SYMMETRIC | I.SPMV_PLAN_NO_DEVICE)
