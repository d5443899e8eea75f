"""GPU parity of the pipelined-PCG vector kernels (sblas) against the CPU oracle."""
import numpy as np
import pytest

import ispm_amd as I
import ispm_amd.gen as G
import oracle as O

pytestmark = pytest.mark.gpu


def dev(a, dt=np.float64):
    return I.DeviceArray.from_numpy(np.asarray(a, dt))


@pytest.mark.parametrize("n", [0, 1, 5, 1000, 1000003])
@pytest.mark.parametrize("t", [I.F64, I.F32])
def test_dot(gpu, n, t):
    dt = np.float64 if t == I.F64 else np.float32
    x = G.uniform(1, n).astype(dt)
    y = G.uniform(2, n).astype(dt)
    s = I.DeviceArray(1, dt)
    for _ in range(3):   # repeated calls reuse the completion counter
        I.dot(t, dev(x, dt), dev(y, dt), s, n)
        got = float(s.numpy()[0])
        ref = O.dot(x.astype(np.float64), y.astype(np.float64))
        scale = float(np.abs(x.astype(np.float64) * y).sum()) + 1e-300
        assert abs(got - ref) <= (1e-14 if t == I.F64 else 1e-7) * scale


def test_dot_unaligned(gpu):
    n = 10001
    x = G.uniform(3, n + 1)
    xd = I.DeviceArray.from_numpy(x)
    s = I.DeviceArray(1)
    import ctypes as C
    off = C.c_void_p(xd.ptr + 8)
    I.check(I.lib().ispm_dot(I.F64, off, off, s.vp, n))
    ref = O.dot(x[1:], x[1:])
    assert abs(s.numpy()[0] - ref) <= 1e-14 * ref


def test_update_scalars(gpu):
    # beta = g/gold; alpha = g / (delta - beta g / alpha_old)
    for t, dt, tol in ((I.F64, np.float64, 1e-15), (I.F32, np.float32, 1e-6)):
        al, be = dev([0.5], dt), dev([0.0], dt)
        g, go, d = dev([3.0], dt), dev([2.0], dt), dev([10.0], dt)
        I.ppcg_update_scalars(t, al, be, g, go, d)
        a_ref, b_ref = O.ppcg_update_scalars(0.5, 0.0, 3.0, 2.0, 10.0)
        assert b_ref == 1.5 and abs(a_ref - 3.0 / (10.0 - 1.5 * 3.0 / 0.5)) < 1e-15
        assert abs(be.numpy()[0] - b_ref) <= tol * abs(b_ref)
        assert abs(al.numpy()[0] - a_ref) <= tol * abs(a_ref)
    # first-iteration convention: gammaold = inf gives beta = 0, alpha = gamma/delta
    al, be = dev([1.0]), dev([0.0])
    I.ppcg_update_scalars(I.F64, al, be, dev([3.0]), dev([np.inf]), dev([4.0]))
    assert be.numpy()[0] == 0.0 and al.numpy()[0] == 0.75


@pytest.mark.parametrize("n", [1, 7, 4099, 300001])
@pytest.mark.parametrize("t", [I.F64, I.F32])
@pytest.mark.parametrize("offset", [0, 1])
def test_update_vectors(gpu, n, t, offset):
    dt = np.float64 if t == I.F64 else np.float32
    vecs = [G.uniform(10 + k, n + offset).astype(dt) for k in range(10)]
    a, b = 0.37, -0.81
    ad, bd = dev([a], dt), dev([b], dt)
    devs = [I.DeviceArray.from_numpy(v) for v in vecs]
    import ctypes as C
    ptr = [C.c_void_p(d.ptr + offset * np.dtype(dt).itemsize) for d in devs]
    rn, g, d = I.DeviceArray(1, dt), I.DeviceArray(1, dt), I.DeviceArray(1, dt)
    I.check(I.lib().ispm_ppcg_update_vectors(t, rn.vp, g.vp, d.vp, ad.vp, bd.vp, *ptr, n))
    got = [dv.numpy()[offset:] for dv in devs]
    nv, mv, p, s, q, z, x, r, u, w = [v[offset:].astype(np.float64) for v in vecs]
    af = float(np.asarray(a, dt)); bf = float(np.asarray(b, dt))
    ref = O.ppcg_update_vectors(af, bf, nv, mv, p, s, q, z, x, r, u, w)
    tol = 1e-14 if t == I.F64 else 1e-5
    names = ["p", "s", "q", "z", "x", "r", "u", "w"]
    for name, gv, rv in zip(names, got[2:], ref[3:]):
        assert np.all(np.abs(gv.astype(np.float64) - rv) <= tol * (np.abs(rv) + 4)), name
    for gs, rs, sc in ((rn, ref[0], ref[0]), (g, ref[1], np.abs(ref[1]) + n), (d, ref[2], np.abs(ref[2]) + n)):
        assert abs(float(gs.numpy()[0]) - rs) <= tol * 10 * sc


def test_indexed(gpu):
    n = 5000
    src = G.uniform(30, n)
    idx = np.random.default_rng(0).permutation(n).astype(np.int32)
    dst = I.DeviceArray(n)
    I.copy_indexed(I.F64, dst, dev(src), I.DeviceArray.from_numpy(idx), n)
    ref = np.zeros(n)
    O.lib().orc_copy_indexed(O.P(ref), O.P(src), O.P(idx, O._ip), n)
    np.testing.assert_array_equal(dst.numpy(), ref)
    base = G.uniform(31, n)
    dd = dev(base)
    I.add_indexed(I.F64, dd, dev(src), I.DeviceArray.from_numpy(idx), n)
    ref = base.copy()
    O.lib().orc_add_indexed(O.P(ref), O.P(src), O.P(idx, O._ip), n)
    np.testing.assert_array_equal(dd.numpy(), ref)
