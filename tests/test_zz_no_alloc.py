"""ispm_ppcg_update_vectors_noalloc (the entry point that allocates nothing: its reduction
partials live in a static device array) against ispm_ppcg_update_vectors, through the C-ABI:
the same outputs bit for bit, on two calls in a row, the second with a smaller input, and
within the oracle's bars (FP64 1e-14, FP32 1e-5 relative)."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

import ispm_amd as I
import ispm_amd.gen as G
import oracle as O

pytestmark = pytest.mark.gpu

NAMES = ["p", "s", "q", "z", "x", "r", "u", "w"]


def run(entry, t, n, seed):
    dt = np.float64 if t == I.F64 else np.float32
    vecs = [G.uniform(seed + k, n).astype(dt) for k in range(10)]
    a, b = 0.37, -0.81
    ad = I.DeviceArray.from_numpy(np.asarray([a], dt))
    bd = I.DeviceArray.from_numpy(np.asarray([b], dt))
    devs = [I.DeviceArray.from_numpy(v) for v in vecs]
    rn, g, d = I.DeviceArray(1, dt), I.DeviceArray(1, dt), I.DeviceArray(1, dt)
    st = entry(t, rn.vp, g.vp, d.vp, ad.vp, bd.vp, *[v.vp for v in devs], n)
    assert st == 0, I.lib().ispm_last_error()
    out = [v.numpy() for v in devs[2:]]
    return vecs, out, [float(x.numpy()[0]) for x in (rn, g, d)]


@pytest.mark.parametrize("t", [I.F64, I.F32], ids=["f64", "f32"])
def test_noalloc_matches_allocating_entry_point(gpu, t):
    L = I.lib()
    # first call at 8192 workgroups (the array's capacity), then a much smaller one: the
    # second must not see the first call's partials
    for n, seed in ((4_500_001, 100), (30_007, 200)):
        vecs, ref_v, ref_s = run(L.ispm_ppcg_update_vectors, t, n, seed)
        _, got_v, got_s = run(L.ispm_ppcg_update_vectors_noalloc, t, n, seed)
        for name, gv, rv in zip(NAMES, got_v, ref_v):
            np.testing.assert_array_equal(gv, rv, err_msg=f"{name}, n={n}")
        assert got_s == ref_s, (n, got_s, ref_s)
        # and the oracle's recurrences (util/cuda/sblas.cu:177-237) within the sblas bars
        dt = vecs[0].dtype
        af, bf = float(np.asarray(0.37, dt)), float(np.asarray(-0.81, dt))
        ref = O.ppcg_update_vectors(af, bf, *[v.astype(np.float64) for v in vecs])
        tol = 1e-14 if t == I.F64 else 1e-5
        for name, gv, rv in zip(NAMES, got_v, ref[3:]):
            assert np.all(np.abs(gv.astype(np.float64) - rv) <= tol * (np.abs(rv) + 4)), name
        for gs, rs, sc in ((got_s[0], ref[0], ref[0]), (got_s[1], ref[1], abs(ref[1]) + n),
                           (got_s[2], ref[2], abs(ref[2]) + n)):
            assert abs(gs - rs) <= tol * 10 * sc


def test_noalloc_rejects_negative_size(gpu):
    L = I.lib()
    z = I.DeviceArray(16)
    st = L.ispm_ppcg_update_vectors_noalloc(I.F64, *([z.vp] * 15), -1)
    assert st == I.ERR_INVALID
    st = L.ispm_ppcg_update_vectors_noalloc(7, *([z.vp] * 15), 4)
    assert st == I.ERR_UNSUPPORTED


def test_noalloc_needs_init_first(gpu):
    """In a process that never initialised the library, the entry point refuses (it would
    have to allocate to initialise) instead of allocating."""
    code = (
        "import ctypes as C, sys\n"
        f"L = C.CDLL({I.LIB_PATH!r})\n"
        "L.ispm_ppcg_update_vectors_noalloc.argtypes = [C.c_int] + [C.c_void_p] * 15 + [C.c_int]\n"
        "p = C.c_void_p(16)\n"
        "st = L.ispm_ppcg_update_vectors_noalloc(1, *([p] * 15), 4)\n"
        "L.ispm_last_error.restype = C.c_char_p\n"
        "print(st, L.ispm_last_error().decode())\n"
    )
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    st, msg = out.stdout.split(" ", 1)
    assert int(st) == I.ERR_INVALID and "ispm_init" in msg
