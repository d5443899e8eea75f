"""The headline workload at its full size: 27-point stencil on 192^3 (7,077,888 rows,
189,119,224 nonzeros), coefficient contrast 1e4 in random 8^3 blocks, approximate-inverse
preconditioner at the default drop tolerance 0.125 (fastainv.h:14 of the reference).

Checked against the CPU oracle:
  * the mixed-precision PCG exactly as it is timed (A in FP32 symmetric diagonal storage, the
    preconditioner as the product Ainv2*Ainv1 in a BF16 split-diagonal plan, FP32 vectors,
    FP64 dot products, FP64 true residual through the FP64-vector/FP32-element product of the
    same stored A) reaches ||b - A x||_2 / ||b||_2 <= 1e-10 by the oracle's FP64 CSR product;
  * the best FP64 solve (Hestenes-Stiefel PCG, FP64 symmetric A, FP64 split product) does too;
  * the BF16 split-diagonal product and BF16 SlELL products of the full-size matrices agree
    per entry with the oracle's product of the same BF16-rounded values (FP32 accumulation:
    |dy| <= 1e-5 sum_k |a_rk x_k|), and the FP64 split product within 1e-13 of that sum.
"""
import numpy as np
import pytest

import ispm_amd as I
import ispm_amd.gen as G
import oracle as O
from ispm_amd import solver as S

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rounded(A, et):
    v = A.vals.astype(np.float64)
    if et == I.F32:
        v = v.astype(np.float32).astype(np.float64)
    elif et == I.BF16:
        f = v.astype(np.float32).view(np.uint32).astype(np.uint64)
        f = ((f + 0x7FFF + ((f >> 16) & 1)) >> 16) << 16      # round to nearest even
        v = f.astype(np.uint32).view(np.float32).astype(np.float64)
    return I.CSR(A.n_rows, A.n_cols, A.row_ptr, A.cols, v)


@pytest.fixture(scope="module")
def work(gpu):
    print("generating 27-point 192^3 contrast matrix", flush=True)
    A = G.stencil27_contrast(192)
    assert A.n_rows == 7077888 and A.nnz == 189119224
    M = I.fastainv_sym(A, I.F64, max_elts=64 * A.n_rows, droptol=0.125)
    P = I.csr_multiply(M[1], M[0], I.F64)
    b = O.csr_spmv(A, G.uniform(1, A.n_rows))
    return {"A": A, "M": M, "P": P, "b": b}


def relres(A, b, x):
    return float(np.linalg.norm(b - O.csr_spmv(A, x)) / np.linalg.norm(b))


def spmv_check(A, plan, vt, et, seed):
    n = plan.n_rows_padded
    dt = np.float64 if vt == I.F64 else np.float32
    x = G.uniform(seed, A.n_cols)
    xd = I.DeviceArray.from_numpy(np.concatenate([x, np.zeros(n - A.n_cols)]).astype(dt))
    yd = I.DeviceArray(n, dt)
    plan.execute(xd, yd)
    y = yd.numpy()
    Ar = rounded(A, et)
    xr = x if vt == I.F64 else x.astype(np.float32).astype(np.float64)
    yref = O.csr_spmv(Ar, xr)
    scale = O.csr_spmv_abs(Ar, xr)
    bar = 1e-13 if vt == I.F64 else 1e-5
    got = y[:A.n_rows].astype(np.float64)
    bad = np.abs(got - yref) > bar * scale + 1e-300
    assert not bad.any(), f"{int(bad.sum())} rows off, worst {np.abs(got - yref)[bad].max():.3e}"
    assert np.all(y[A.n_rows:] == 0)


@pytest.mark.parametrize("vt,et", [(I.F32, I.BF16), (I.F64, I.F64)], ids=["f32-bf16", "f64-f64"])
def test_split_product_full_size(work, vt, et):
    """The preconditioner product Ainv2*Ainv1 (1.29 nonzeros per row) in a split-diagonal plan."""
    P = work["P"]
    plan = I.Plan(P, vt, elm_type=et, flags=I.SPMV_PLAN_SPLIT_DIAG)
    spmv_check(P, plan, vt, et, 11)


def test_bf16_slell_full_size(work):
    """BF16 elements in SlELL layouts: the preconditioner product in the fewest-bytes format
    (streamed kernel) and A in the default 27-diagonal format (shared-pattern kernel)."""
    P, A = work["P"], work["A"]
    spmv_check(P, I.Plan(P, I.F32, elm_type=I.BF16, kind=I.PLAN_MIN_BYTES), I.F32, I.BF16, 12)
    pa = I.Plan(A, I.F32, elm_type=I.BF16)
    assert pa.info().pattern_planes == 27
    spmv_check(A, pa, I.F32, I.BF16, 13)


def test_guarded_symmetric_full_size(work):
    """The FP32 product of A the mixed solve runs (x with zero guard zones, four rows per
    lane): per entry against the oracle, and bit-identical to the checked kernel."""
    A = work["A"]
    p = I.Plan(A, I.F32, elm_type=I.F32, flags=I.SPMV_PLAN_SYMMETRIC)
    g = p.x_guard()
    n = p.n_rows_padded
    x = G.uniform(14, A.n_rows)
    xg = np.zeros(n + 2 * g, np.float32)
    xg[g:g + A.n_rows] = x
    yd = I.DeviceArray(n, np.float32)
    p.execute_guarded(I.DeviceArray.from_numpy(xg), g, yd)
    y = yd.numpy()
    Ar = rounded(A, I.F32)
    xr = x.astype(np.float32).astype(np.float64)
    yref = O.csr_spmv(Ar, xr)
    bad = np.abs(y[:A.n_rows].astype(np.float64) - yref) > 1e-5 * O.csr_spmv_abs(Ar, xr)
    assert not bad.any(), int(bad.sum())
    assert np.all(y[A.n_rows:] == 0)
    x2 = I.DeviceArray.from_numpy(np.concatenate([x, np.zeros(n - A.n_rows)]).astype(np.float32))
    y2 = I.DeviceArray(n, np.float32)
    p.execute(x2, y2)
    np.testing.assert_array_equal(y, y2.numpy())


def headline_solver(work, precision):
    s = S.Solver(work["A"], M=work["M"], precision=precision, method=S.PCG, tol=TOL,
                 symmetric=True, fsai_product=True, split_diag=True)
    assert s.A_plan.flags & I.SPMV_PLAN_SYMMETRIC
    assert len(s.M_plans) == 1 and s.M_plans[0].flags & I.SPMV_PLAN_SPLIT_DIAG
    # the coefficients (1, 1e4 and their half sums) are exact in FP32: the FP64 residual uses
    # the FP64-vector product of the same FP32 storage, no second plan
    assert s.res_plan is None
    return s


def solve(s, b):
    n = s.n
    bd = I.DeviceArray.from_numpy(np.concatenate([b, np.zeros(n - b.size)]))
    xd = I.DeviceArray(n)
    st = s.solve_device(bd, xd, allow_fail=True)
    return st, xd.numpy()[:b.size]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_headline_solve_full_size(work, precision):
    A, b = work["A"], work["b"]
    s = headline_solver(work, precision)
    st, x = solve(s, b)
    rel = relres(A, b, x)
    print(f"{precision}: {st.iterations} iterations, {st.replacements} replacements, "
          f"oracle relres {rel:.3e}", flush=True)
    assert st.converged
    assert rel <= TOL
    work.setdefault("iterations", {})[precision] = st.iterations
    if precision == "mixed":
        assert st.replacements >= 1
    # the FP32 inner iteration costs (almost) no iterations against the FP64 solve
    if len(work["iterations"]) == 2:
        im, i64 = work["iterations"]["mixed"], work["iterations"]["fp64"]
        assert abs(im - i64) <= max(10, i64 // 20), (im, i64)
