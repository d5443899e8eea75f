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


def _compile(tmp_path, body):
    import subprocess
    src = tmp_path / "tu.cpp"
    src.write_text(body)
    pkg = I.PKG_DIR
    return subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", pkg, "-I", os.path.join(ROOT, "include"),
                           str(src)], capture_output=True, text=True)


def test_reference_names_compile(tmp_path):
    """A translation unit naming the reference's allocators, timer, plan kinds and pinned
    helpers (util/cuda/allocator.h:8,45, timer.h:6,33, plan/plan.h:12-16, pinner.h:7-48)
    compiles against the drop-in headers."""
    r = _compile(tmp_path, """
#include <vector>
#include "plan/plan.h"
#include "util/cuda/allocator.h"
#include "util/cuda/pinner.h"
#include "util/cuda/timer.h"
void f(const csr_matrix<double> &m) {
  std::vector<double, cuda_device_allocator<double> > d(4);
  std::vector<double, cuda_pinned_allocator<double> > h(4);
  cudatimer t; t.start(); t.stop(); (void)t.elapsed_seconds();
  cudaEventScoped e; e.record(); e.await();
  spmv_plan<double> p(m, PLAN_ESTIMATE);
  pinmemptr<double> pm(h.data()); (void)pm();
  cuda_pinned_region<double> unused;
}
""")
    assert r.returncode == 0, r.stderr


def test_pinned_region_registration_is_a_compile_error(tmp_path):
    """Registering existing host memory is not available on this target: constructing a
    cuda_pinned_region from a pointer fails at compile time with a message saying so."""
    r = _compile(tmp_path, """
#include "util/cuda/pinner.h"
void f(double *p) { cuda_pinned_region<double> r(p, 4); }
""")
    assert r.returncode != 0
    assert "registering existing host memory is not supported" in r.stderr
