"""A caller written against the reference's C++ API (tests/cxx/api_smoke.cpp, built by the
library Makefile against the drop-in headers) runs on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cxx_api_smoke(gpu):
    exe = os.path.join(ROOT, "tests", "cxx", "api_smoke.bin")
    assert os.path.exists(exe), "build the library first (make -C ispm-sparse-lib_amd)"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stderr
    assert "ok all" in out.stdout


def test_cxx_foam_call_pattern(gpu):
    """The foam-extend call pattern through the C-ABI alone (tests/cxx/foam_smoke.cpp):
    lduMatrix to CSR, FSAI-PCG with OpenFOAM's normalised residual (checked against the
    formula on the host), and a second time step refreshed on the device."""
    exe = os.path.join(ROOT, "tests", "cxx", "foam_smoke.bin")
    assert os.path.exists(exe), "build the library first (make -C ispm-sparse-lib_amd)"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stderr
    assert out.stdout.count("ok step") == 2 and "ok all" in out.stdout
