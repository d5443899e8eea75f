"""The foam-extend-shaped driver on the GPU: lduMatrix coefficients scattered on the device,
OpenFOAM's normalised residual (checked with the oracle's restatement), tolerance and relTol
stops, an initial field that already meets the tolerance, plan reuse across time steps with
the coefficient refresh, drop-tolerance probing across steps, and mixed precision."""
import numpy as np
import pytest

import ispm_amd as I
import ispm_amd.gen as G
from ispm_amd import foam as F
import oracle as O

pytestmark = pytest.mark.gpu


def _ldu(m=20, seed=4001, sigma=1.0):
    return F.LduMatrix(*G.ldu_mesh(m, m, m, sigma=sigma, seed=seed))


def _system(L, seed=5):
    A, _ = F.ldu_to_csr(L)
    b = O.csr_spmv(A, G.uniform(seed, L.n))
    return A, b


def test_ldu_values_device_matches_host(gpu):
    L = F.LduMatrix(*G.ldu_mesh(9, 7, 5, seed=11, asymmetry=0.4))
    A, pos = F.ldu_to_csr(L)
    d = [I.DeviceArray.from_numpy(a) for a in (pos, L.diag, L.upper, L.lower)]
    vals = I.DeviceArray(A.nnz, np.float64)
    I.check(I.lib().ispm_ldu_values_device(L.n, L.n_faces, d[0].vp, d[1].vp, d[2].vp, d[3].vp, vals.vp))
    np.testing.assert_array_equal(vals.numpy(), np.asarray(A.vals))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_normalised_tolerance(gpu, precision):
    L = _ldu()
    A, b = _system(L)
    psi0 = G.uniform(9, L.n)
    drv = F.FoamPCG(L, precision=precision, droptol=0.125)
    psi, perf = drv.solve(b, psi0, tolerance=1e-8)
    assert perf.converged and perf.n_iterations > 0
    nf, r = O.foam_residual(A, b, psi0, psi)
    _, r0 = O.foam_residual(A, b, psi0, psi0)
    assert r < 1e-8
    assert abs(perf.final_residual - r) <= 1e-6 * r + 1e-18
    assert abs(perf.initial_residual - r0) <= 1e-9 * r0


def test_normalised_rel_tol(gpu):
    L = _ldu(seed=4002)
    A, b = _system(L, seed=6)
    psi0 = np.zeros(L.n)
    drv = F.FoamPCG(L, precision="fp64", droptol=0.125)
    psi, perf = drv.solve(b, psi0, tolerance=1e-14, rel_tol=1e-4)
    _, r = O.foam_residual(A, b, psi0, psi)
    _, r0 = O.foam_residual(A, b, psi0, psi0)
    assert perf.converged and r < 1e-4 * r0
    _, tight = drv.solve(b, psi0, tolerance=1e-10)
    assert tight.n_iterations > perf.n_iterations


def test_initial_field_already_converged(gpu):
    L = _ldu(m=12, seed=4003)
    A, b = _system(L, seed=7)
    drv = F.FoamPCG(L, precision="fp64", droptol=0.125)
    psi, perf = drv.solve(b, np.zeros(L.n), tolerance=1e-12)
    assert perf.converged
    psi2, perf2 = drv.solve(b, psi, tolerance=1e-6)
    assert perf2.converged and perf2.n_iterations == 0
    np.testing.assert_array_equal(psi2, psi)
    assert perf2.final_residual == perf2.initial_residual < 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("symmetric", [True, False], ids=["symmetric", "slell"])
def test_time_steps_reuse_plan_and_probe_droptol(gpu, symmetric, precision):
    """Three time steps on one addressing: the device refresh of the A plan (symmetric
    diagonal storage or SlELL) and of the mixed solve's FP64 residual plan, the solves to the
    normalised tolerance by the oracle's measure, and the tuner's first probes."""
    L = _ldu(m=16, seed=4010)
    drv = F.FoamPCG(L, precision=precision, symmetric=symmetric)
    assert bool(drv.A_plan.flags & I.SPMV_PLAN_SYMMETRIC) == symmetric
    flags = I.SPMV_PLAN_SYMMETRIC if symmetric else 0
    vt = I.F32 if precision == "mixed" else I.F64
    psi = np.zeros(L.n)
    probes = []
    for step in range(3):
        Ls = _ldu(m=16, seed=4010 + step, sigma=1.0 + 0.2 * step)   # same addressing
        if step:
            drv.update(Ls)
        A, b = _system(Ls, seed=20 + step)
        psi0 = psi.copy()
        psi, perf = drv.solve(b, psi0, tolerance=1e-9)
        _, r = O.foam_residual(A, b, psi0, psi)
        assert perf.converged and r < 1e-9
        probes.append(perf.droptol)
        # the preconditioner storage chosen at the first solve is kept for the later ones
        if step == 0:
            storage = drv.storage
            assert storage is not None
        assert drv.storage == storage
        # the refreshed plans multiply exactly like plans built from this step's matrix
        n = drv.n_pad
        dt = np.float32 if vt == I.F32 else np.float64
        x = I.DeviceArray.from_numpy(np.concatenate([G.uniform(30 + step, L.n), np.zeros(n - L.n)]).astype(dt))
        y1, y2 = I.DeviceArray(n, dt), I.DeviceArray(n, dt)
        drv.A_plan.execute(x, y1)
        I.Plan(A, vt, elm_type=vt, flags=flags).execute(x, y2)
        np.testing.assert_array_equal(y1.numpy(), y2.numpy())
        if drv.res_plan is not None:
            x64 = I.DeviceArray.from_numpy(np.concatenate([G.uniform(40 + step, L.n), np.zeros(n - L.n)]))
            y3, y4 = I.DeviceArray(n), I.DeviceArray(n)
            drv.res_plan.execute(x64, y3)
            I.Plan(A, I.F64, elm_type=I.F64, flags=flags).execute(x64, y4)
            np.testing.assert_array_equal(y3.numpy(), y4.numpy())
    # the first two probes of the golden-section search on [2^-10, 2^-3]
    assert abs(np.log2(probes[0]) - (-10 + 7 * (3 - 5 ** 0.5) / 2)) < 1e-12
    assert probes[1] != probes[0]


def test_update_rejects_other_addressing(gpu):
    drv = F.FoamPCG(_ldu(m=8), precision="fp64", droptol=0.125)
    with pytest.raises(ValueError):
        drv.update(_ldu(m=9))
