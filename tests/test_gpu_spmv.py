"""GPU parity of the SlELL SpMV / SpMM against the CPU oracle (CSR product in FP64).

Bars: integer-valued A and x -> bit-exact; FP64 -> |dy| <= 1e-13 sum_k |a_rk x_k| per entry;
FP32 vectors -> |dy| <= 1e-5 sum_k |a_rk x_k| (bf16/fp32 element values are compared against
the oracle run on the same rounded values, so only the accumulation differs).
"""
import numpy as np
import pytest

import ispm_amd as I
import ispm_amd.gen as G
import oracle as O

pytestmark = pytest.mark.gpu

FORMATS = [dict(S=32), dict(S=64), dict(S=256), dict(S=64, D=True), dict(S=256, D=True),
           dict(S=128, hblock=2), dict(S=32, hblock=2), dict(S=256, V=True),
           dict(S=64, V=True, D=True), dict(S=128, V=True, hblock=2), dict(S=512),
           dict(S=4), dict(S=8, D=True), dict(S=16, hblock=2)]
LAUNCH = [(64, 1), (128, 1), (256, 2), (512, 1), (1024, 2), (64, 2)]
TYPES = [(I.F64, I.F64), (I.F64, I.F32), (I.F32, I.F32), (I.F32, I.BF16)]


def rounded_vals(A, et):
    v = A.vals.astype(np.float64)
    if et == I.F32:
        return v.astype(np.float32).astype(np.float64)
    if et == I.BF16:
        f = v.astype(np.float32).view(np.uint32).astype(np.uint64)
        f = ((f + 0x7FFF + ((f >> 16) & 1)) >> 16) << 16
        return f.astype(np.uint32).view(np.float32).astype(np.float64)
    return v


def run_spmv(hs, x, vt, N, H, mode=0):
    dm = I.DeviceSlell(hs)
    dt = np.float64 if vt == I.F64 else np.float32
    xv = np.zeros(max(hs.n_rows, hs.n_cols), dt)
    xv[:x.size] = x
    xd = I.DeviceArray.from_numpy(xv)
    yd = I.DeviceArray(hs.n_rows, dt)
    yd.upload(np.full(hs.n_rows, 7.0, dt))
    st = dm.spmv(xd, yd, vt, N=N, S=hs.slice_height if not hs.var_height else N * H, H=H, mode=mode)
    I.synchronize()
    return st, yd.numpy()


def check(A, hs, x, y, vt, et, integer=False):
    Ar = I.CSR(A.n_rows, A.n_cols, A.row_ptr, A.cols, rounded_vals(A, et))
    xr = x.astype(np.float64) if vt == I.F64 else x.astype(np.float32).astype(np.float64)
    yref = O.csr_spmv(Ar, xr)
    scale = O.csr_spmv_abs(Ar, xr)
    got = y[:A.n_rows].astype(np.float64)
    if integer:
        np.testing.assert_array_equal(got, yref)
    else:
        tol = 1e-13 if vt == I.F64 else 1e-5
        bad = np.abs(got - yref) > tol * scale + 1e-300
        assert not bad.any(), f"{bad.sum()} rows off; worst {np.abs(got - yref)[bad].max()}"
    # padded rows are written as zero
    assert np.all(y[A.n_rows:] == 0)


@pytest.fixture(scope="module")
def mats(gpu):
    return {
        "poisson7": G.poisson7(16),
        "irregular": G.random_sparse(3000, 11, min_len=1, max_len=60),
        "pairs": G.random_sparse(2500, 12, min_len=2, max_len=30, pair_bias=0.7),
        "skewed": G.random_sparse(1800, 13, min_len=0, max_len=200),
        "short_long": short_long_rows(),
    }


def short_long_rows(n=6000, seed=14):
    """Rows of 1-3 entries with every 700th row holding 2500: workgroups of the streamed
    mapping (256 rows) fall on both sides of its LDS capacity."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(1, 4, n)
    lens[::700] = 2500
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    cols = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens]).astype(np.int32)
    vals = rng.uniform(-1, 1, rp[-1])
    return I.CSR(n, n, rp, cols, vals)


def stream_ok(hs):
    return not hs.var_height and hs.hblock == 1 and not hs.diags and hs.slice_height <= 256


@pytest.mark.parametrize("fmt", FORMATS, ids=lambda f: "-".join(f"{k}{int(v)}" for k, v in f.items()))
@pytest.mark.parametrize("name", ["poisson7", "irregular", "pairs", "skewed"])
def test_spmv_formats_fp64(mats, name, fmt):
    A = mats[name]
    hs = I.csr_to_slell(A, fmt["S"], fmt.get("hblock", 1), fmt.get("V", False), fmt.get("D", False))
    x = G.uniform(100, A.n_cols)
    ran = 0
    for N, H in LAUNCH:
        S = hs.slice_height if not hs.var_height else N * H
        if not (N >= S // H and S >= H):
            continue
        st, y = run_spmv(hs, x, I.F64, N, H)
        assert st == 0, I.lib().ispm_last_error()
        check(A, hs, x, y, I.F64, I.F64)
        ran += 1
    if not hs.var_height and hs.slice_height < 64:
        st, y = run_spmv(hs, x, I.F64, 256, 1, mode=1)      # packed rows
        assert st == 0, I.lib().ispm_last_error()
        check(A, hs, x, y, I.F64, I.F64)
    if not hs.var_height and hs.slice_height <= 256:
        # streamed rows (plain planes), else the packed / slice-per-workgroup fallback
        st, y = run_spmv(hs, x, I.F64, 256, 1, mode=2)
        assert st == 0, I.lib().ispm_last_error()
        check(A, hs, x, y, I.F64, I.F64)
    assert ran > 0


def stencil_then_irregular():
    """27-point rows (16^3 grid) followed by irregular rows in their own column block:
    slices with the shared diagonal pattern and slices without it."""
    A = G.stencil27_contrast(16)
    B = G.random_sparse(4096, 15, min_len=1, max_len=30)
    rp = np.concatenate([A.row_ptr, A.row_ptr[-1] + B.row_ptr[1:]]).astype(np.int64)
    cols = np.concatenate([A.cols, B.cols + A.n_cols]).astype(np.int32)
    vals = np.concatenate([A.vals, B.vals])
    return I.CSR(A.n_rows + B.n_rows, A.n_cols + B.n_cols, rp, cols, vals)


@pytest.mark.parametrize("S,H", [(64, 1), (128, 1), (256, 1), (128, 2), (256, 2), (512, 1), (512, 2)])
@pytest.mark.parametrize("vt,et", TYPES)
def test_spmv_diag_pattern(mats, S, H, vt, et):
    """Plans whose diagonal slices mostly share one pattern: those slices stage the x windows
    of the pattern in LDS, the others read x directly (one kernel, both paths)."""
    for name, A, planes in (("poisson7", mats["poisson7"], 7), ("stencil27", G.stencil27_contrast(16), 27),
                            ("mixed", stencil_then_irregular(), 27)):
        p = I.Plan(A, vt, elm_type=et, explicit=dict(S=S, D=True), launch=I.Launch(S // H, S, H, 0, 0, 0))
        inf = p.info()
        assert inf.pattern_planes == planes, name
        n_slices = p.n_rows_padded // S
        assert 0 < inf.pattern_slices <= n_slices
        if name == "mixed":
            assert inf.pattern_slices < n_slices
        n = p.n_rows_padded
        dt = np.float64 if vt == I.F64 else np.float32
        x = G.uniform(104, A.n_cols)
        xd = I.DeviceArray.from_numpy(np.concatenate([x, np.zeros(n - A.n_cols)]).astype(dt))
        yd = I.DeviceArray(n, dt)
        yd.upload(np.full(n, 7.0, dt))
        p.execute(xd, yd)
        I.synchronize()
        check(A, None, x, yd.numpy(), vt, et)


@pytest.mark.parametrize("S", [1, 2, 4, 8, 32, 256])
@pytest.mark.parametrize("vt,et", TYPES)
def test_spmv_stream_capacity(mats, S, vt, et):
    """Streamed mapping on rows that put some workgroups over the LDS capacity."""
    A = mats["short_long"]
    hs = I.csr_to_slell(A, S, out_type=et)
    assert stream_ok(hs)
    x = G.uniform(103, A.n_cols)
    st, y = run_spmv(hs, x, vt, 256, 1, mode=2)
    assert st == 0, I.lib().ispm_last_error()
    check(A, hs, x, y, vt, et)
    if S < 64:
        st, y = run_spmv(hs, x, vt, 256, 1, mode=1)
        assert st == 0, I.lib().ispm_last_error()
        check(A, hs, x, y, vt, et)


@pytest.mark.parametrize("vt,et", TYPES)
@pytest.mark.parametrize("fmt", [dict(S=64, D=True), dict(S=128, hblock=2), dict(S=256, V=True), dict(S=32),
                                 dict(S=8, mode=1), dict(S=4, D=True, mode=1), dict(S=4, mode=2),
                                 dict(S=16, mode=2)],
                         ids=lambda f: "-".join(f"{k}{int(v)}" for k, v in f.items()))
def test_spmv_types(mats, vt, et, fmt):
    A = mats["pairs"] if fmt.get("hblock") else mats["irregular"]
    hs = I.csr_to_slell(A, fmt["S"], fmt.get("hblock", 1), fmt.get("V", False), fmt.get("D", False), out_type=et)
    x = G.uniform(101, A.n_cols)
    N, H = (256, 1)
    st, y = run_spmv(hs, x, vt, N, H, mode=fmt.get("mode", 0))
    assert st == 0, I.lib().ispm_last_error()
    check(A, hs, x, y, vt, et)


@pytest.mark.parametrize("fmt", FORMATS[:8], ids=lambda f: "-".join(f"{k}{int(v)}" for k, v in f.items()))
def test_spmv_integer_exact(mats, fmt):
    A0 = G.random_sparse(2000, 21, min_len=0, max_len=50, integer=True, pair_bias=0.5)
    hs = I.csr_to_slell(A0, fmt["S"], fmt.get("hblock", 1), fmt.get("V", False), fmt.get("D", False))
    x = np.round(G.uniform(102, A0.n_cols) * 16)
    for N, H in ((64, 1), (256, 2), (1024, 1)):
        S = hs.slice_height if not hs.var_height else N * H
        if not (N >= S // H and S >= H):
            continue
        st, y = run_spmv(hs, x, I.F64, N, H)
        assert st == 0
        check(A0, hs, x, y, I.F64, I.F64, integer=True)
    if stream_ok(hs):
        st, y = run_spmv(hs, x, I.F64, 256, 1, mode=2)
        assert st == 0
        check(A0, hs, x, y, I.F64, I.F64, integer=True)


def test_invalid_launch_is_noop(mats):
    A = mats["poisson7"]
    hs = I.csr_to_slell(A, 512)
    st, y = run_spmv(hs, G.uniform(1, A.n_cols), I.F64, 64, 1)   # N < S/H: reference no-op
    assert st == 2
    assert np.all(y == 7.0)


def test_spmm_matches_spmv(mats):
    A = mats["irregular"]
    for fmt in (dict(S=64, D=True), dict(S=128), dict(S=256, V=True)):
        p = I.Plan(A, I.F64, explicit=fmt)
        n = p.n_rows_padded
        for k in (3, 8):
            X = np.zeros((k, n))
            X[:, :A.n_cols] = G.uniform(200 + k, k * A.n_cols).reshape(k, A.n_cols)
            Xd = I.DeviceArray.from_numpy(X.ravel())
            Yd = I.DeviceArray(k * n)
            p.execute_multi(k, Xd, n, Yd, n)
            Y = Yd.numpy().reshape(k, n)
            for j in range(k):
                xd = I.DeviceArray.from_numpy(X[j])
                yd = I.DeviceArray(n)
                p.execute(xd, yd)
                np.testing.assert_array_equal(Y[j], yd.numpy())


def off_pattern(A, rows, seed=17):
    """A with one extra entry (far from the diagonal) in each of the given rows: the slices
    holding them leave the shared diagonal pattern."""
    rng = np.random.default_rng(seed)
    rp, cols, vals = np.asarray(A.row_ptr), np.asarray(A.cols), np.asarray(A.vals, np.float64)
    out_c, out_v, out_p = [], [], [0]
    for i in range(A.n_rows):
        c, v = cols[rp[i]:rp[i + 1]], vals[rp[i]:rp[i + 1]]
        if i in rows:
            extra = int(rng.integers(0, A.n_cols))
            if extra not in c:
                order = np.argsort(np.append(c, extra))
                c, v = np.append(c, extra)[order], np.append(v, 0.25)[order]
        out_c.append(c); out_v.append(v); out_p.append(out_p[-1] + len(c))
    return I.CSR(A.n_rows, A.n_cols, np.array(out_p, np.int64), np.concatenate(out_c).astype(np.int32),
                 np.concatenate(out_v))


@pytest.mark.parametrize("vt,et", [(I.F64, I.F64), (I.F32, I.F32), (I.F64, I.F32)])
@pytest.mark.parametrize("name", ["27pt16", "7pt32", "19pt24_off", "fsaiprod7pt32", "wide27"])
def test_pattern_spmm_matches_spmv(gpu, name, vt, et):
    """Multi-vector products of 256-row diagonal plans with a shared pattern (a block
    solver's stencil A or preconditioner product): every column bit-identical to the single-vector product of
    the same plan, for aligned and unaligned column strides (the latter stage x without
    16-byte loads), with slices off the pattern (19pt24_off) taking the per-row path; and
    within the oracle's bound."""
    if name == "27pt16":
        A = G.stencil27_contrast(16, block=4)
    elif name == "7pt32":
        A = G.poisson7(32)
    elif name == "wide27":   # FP64 windows beyond four vectors per thread
        A = G.stencil(300, 8, 6, 27, seed=8)
    elif name == "19pt24_off":
        A = off_pattern(G.stencil(24, 24, 24, points=19), {5, 700, 701, 9000, 13000})
    else:   # the approximate inverse's product Ainv2 Ainv1 (13 diagonals on most slices)
        a1, a2 = I.fastainv_sym(G.poisson7(32), I.F64, max_elts=64 * 32 ** 3, droptol=0.05)
        A = I.csr_multiply(a2, a1, I.F64)
    dt = np.float64 if vt == I.F64 else np.float32
    p = I.Plan(A, vt, elm_type=et, explicit=dict(S=256, D=True))
    inf = p.info()
    assert inf.pattern_planes > 0 and inf.pattern_slices > 0
    if name == "19pt24_off":
        assert inf.pattern_slices < inf.device.n_slices
    n = p.n_rows_padded
    for k in (2, 3, 8):
        for ld in (n, n + 1):
            X = np.zeros((k, ld), dt)
            X[:, :A.n_cols] = G.uniform(500 + k, k * A.n_cols).reshape(k, A.n_cols)
            Xd = I.DeviceArray.from_numpy(X.ravel())
            Yd = I.DeviceArray(k * ld, dt)
            p.execute_multi(k, Xd, ld, Yd, ld)
            Y = Yd.numpy().reshape(k, ld)
            for j in range(k):
                xd = I.DeviceArray.from_numpy(np.ascontiguousarray(X[j, :n]))
                yd = I.DeviceArray(n, dt)
                p.execute(xd, yd)
                np.testing.assert_array_equal(Y[j, :n], yd.numpy(), err_msg=f"k={k} ld={ld} column {j}")
            hs = I.csr_to_slell(A, 256, 1, False, True, out_type=et)
            check(A, hs, X[k - 1, :A.n_cols], Y[k - 1, :n], vt, et)


@pytest.mark.parametrize("name", ["irregular", "skewed", "short_long"])
@pytest.mark.parametrize("S", [1, 4, 32])
@pytest.mark.parametrize("vt,et", TYPES)
def test_spmm_streamed_matches_spmv(mats, name, S, vt, et):
    """Multi-vector products of plain small-slice layouts in the streamed mapping (the
    block solver's preconditioner factors): every column bit-identical to the single-vector
    streamed product, with and without the plan's layout facts (short_long has groups
    beyond the default LDS capacity)."""
    A = mats[name]
    dt = np.float64 if vt == I.F64 else np.float32
    hs = I.csr_to_slell(A, S, 1, False, False, out_type=et)
    dm = I.DeviceSlell(hs)
    p = I.Plan(A, vt, elm_type=et, explicit=dict(S=S), launch=I.Launch(256, S, 1, 0, 0, 2))
    n = hs.n_rows
    for k in (2, 5, 8):
        X = np.zeros((k, n), dt)
        X[:, :A.n_cols] = G.uniform(400 + k, k * A.n_cols).reshape(k, A.n_cols)
        Xd = I.DeviceArray.from_numpy(X.ravel())
        for how in ("desc", "plan"):
            Yd = I.DeviceArray(k * n, dt)
            if how == "desc":
                assert dm.spmm(k, Xd, n, Yd, n, vt, N=256, mode=2) == 0, I.lib().ispm_last_error()
            else:
                p.execute_multi(k, Xd, n, Yd, n)
            Y = Yd.numpy().reshape(k, n)
            for j in range(k):
                xd = I.DeviceArray.from_numpy(X[j])
                yd = I.DeviceArray(n, dt)
                assert dm.spmv(xd, yd, vt, N=256, mode=2) == 0
                np.testing.assert_array_equal(Y[j], yd.numpy())
            check(A, hs, X[0, :A.n_cols], Y[0], vt, et)


def mostly_diagonal(n=5000, seed=16):
    """Rows holding only their diagonal entry except every 23rd row (1-40 off-diagonal entries
    on both sides); row 7 has no diagonal entry and row 11 a zero one."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        c = {i} if i != 7 else set()
        if i % 23 == 0:
            c |= set(rng.choice(n, rng.integers(1, 41), replace=False).tolist()) - {i}
        rows.append(sorted(c))
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    cols = np.concatenate([np.array(r, np.int32) for r in rows])
    vals = rng.uniform(-1, 1, rp[-1])
    vals[rp[11]:rp[12]][np.array(rows[11]) == 11] = 0.0
    return I.CSR(n, n, rp, cols, vals)


def run_plan(p, A, x, vt):
    n = p.n_rows_padded
    dt = np.float64 if vt == I.F64 else np.float32
    xd = I.DeviceArray.from_numpy(np.concatenate([x, np.zeros(n - A.n_rows)]).astype(dt))
    yd = I.DeviceArray(n, dt)
    yd.upload(np.full(n, 7.0, dt))
    p.execute(xd, yd)
    return yd.numpy()


@pytest.mark.parametrize("vt,et", TYPES)
@pytest.mark.parametrize("name", ["mostly_diagonal", "diagonal", "poisson7", "irregular", "short_long"])
def test_split_diagonal_plan(mats, name, vt, et):
    """Split-diagonal plans (diagonal vector + compressed remainder rows with a row map) give
    the CSR product: the remainder's rows add onto the diagonal part at their matrix rows,
    padded rows stay zero."""
    if name == "mostly_diagonal":
        A = mostly_diagonal()
    elif name == "diagonal":
        A = I.CSR(3000, 3000, np.arange(3001), np.arange(3000, dtype=np.int32), G.uniform(17, 3000))
    else:
        A = mats[name]
    p = I.Plan(A, vt, elm_type=et, flags=I.SPMV_PLAN_SPLIT_DIAG)
    assert p.n_rows_padded % 512 == 0 and p.n_rows_padded >= A.n_rows
    x = G.uniform(18, A.n_rows)
    check(A, None, x, run_plan(p, A, x, vt), vt, et)


def test_split_diagonal_integer_exact(mats):
    A = mostly_diagonal()
    Ai = I.CSR(A.n_rows, A.n_cols, A.row_ptr, A.cols, np.round(A.vals * 8))
    x = np.round(G.uniform(19, A.n_rows) * 8)
    p = I.Plan(Ai, I.F64, flags=I.SPMV_PLAN_SPLIT_DIAG)
    check(Ai, None, x, run_plan(p, Ai, x, I.F64), I.F64, I.F64, integer=True)


def test_split_diagonal_refuses_spmm_and_refresh(mats):
    A = mostly_diagonal()
    p = I.Plan(A, I.F64, flags=I.SPMV_PLAN_SPLIT_DIAG)
    with pytest.raises(I.IspmError):
        p.update_coefs(A)
    n = p.n_rows_padded
    X = I.DeviceArray(2 * n)
    with pytest.raises(I.IspmError):
        p.execute_multi(2, X, n, X, n)


def _sym_matrix(name):
    if name == "poisson7":
        return G.poisson7(16)
    if name == "stencil27_contrast":
        return G.stencil27_contrast(24)
    if name == "stencil19_aniso":
        return G.stencil(20, 18, 16, 19, anisotropy=10, contrast=100, seed=3)
    if name == "levels16":
        # widest offset 4225 spans 16 slices: level order, two positions per XCD, ten
        # slices past the last full level
        return G.stencil(64, 65, 40, 27, seed=6)
    if name == "wide27":
        # 27-point on 300-wide lines: the FP64 window list exceeds four vectors per thread
        return G.stencil(300, 8, 6, 27, seed=8)
    if name == "levels131":
        # widest offset 33668 spans 131 slices: level order with bands of 16 and 17
        # positions, one slice past the last full level
        return G.stencil(131, 256, 3, 27, seed=7)
    return G.stencil(7, 9, 11, 27, seed=4)      # 693 rows: one padded slice, clipped windows


@pytest.mark.parametrize("vt,et", TYPES)
@pytest.mark.parametrize("name", ["poisson7", "stencil27_contrast", "stencil19_aniso", "small27",
                                  "levels16", "levels131"])
def test_symmetric_plan(gpu, name, vt, et):
    """Symmetric diagonal plans (upper planes stored, lower entries read from their mirror
    rows) give the CSR product; padded rows stay zero."""
    A = _sym_matrix(name)
    p = I.Plan(A, vt, elm_type=et, flags=I.SPMV_PLAN_SYMMETRIC)
    inf = p.info()
    assert inf.S == 256 and inf.D == 1 and p.n_rows_padded % 512 == 0
    x = G.uniform(31, A.n_rows)
    check(A, None, x, run_plan(p, A, x, vt), vt, et)


@pytest.mark.parametrize("vt,et", [(I.F64, I.F64), (I.F64, I.F32), (I.F32, I.F32)])
@pytest.mark.parametrize("name", ["poisson7", "stencil27_contrast", "small27", "levels16", "levels131", "wide27"])
def test_symmetric_spmm_matches_spmv(gpu, name, vt, et):
    """Multi-vector products of symmetric diagonal plans: every column bit-identical to the
    plan's single-vector product, for aligned and unaligned column strides, and within the
    oracle's bound; BF16 elements multiply one vector at a time."""
    A = _sym_matrix(name)
    dt = np.float64 if vt == I.F64 else np.float32
    p = I.Plan(A, vt, elm_type=et, flags=I.SPMV_PLAN_SYMMETRIC)
    n = p.n_rows_padded
    for k in (2, 5, 8):
        for ld in (n, n + 1):
            X = np.zeros((k, ld), dt)
            X[:, :A.n_cols] = G.uniform(600 + k, k * A.n_cols).reshape(k, A.n_cols)
            Xd = I.DeviceArray.from_numpy(X.ravel())
            Yd = I.DeviceArray(k * ld, dt)
            p.execute_multi(k, Xd, ld, Yd, ld)
            Y = Yd.numpy().reshape(k, ld)
            for j in range(k):
                xd = I.DeviceArray.from_numpy(np.ascontiguousarray(X[j, :n]))
                yd = I.DeviceArray(n, dt)
                p.execute(xd, yd)
                np.testing.assert_array_equal(Y[j, :n], yd.numpy(), err_msg=f"k={k} ld={ld} column {j}")
            check(A, None, X[k - 1, :A.n_cols], Y[k - 1, :n], vt, et)
    if name == "poisson7":
        pb = I.Plan(A, I.F32, elm_type=I.BF16, flags=I.SPMV_PLAN_SYMMETRIC)
        Xd = I.DeviceArray(2 * pb.n_rows_padded, np.float32)
        with pytest.raises(I.IspmError):
            pb.execute_multi(2, Xd, pb.n_rows_padded, Xd, pb.n_rows_padded)


def _sym_revalued(A):
    """A's pattern with new symmetric values: entry (i, j) scaled by a function of
    (min(i, j), max(i, j))."""
    rows = np.repeat(np.arange(A.n_rows), np.diff(A.row_ptr))
    lo, hi = np.minimum(rows, A.cols), np.maximum(rows, A.cols)
    w = 1.0 + 0.25 * np.cos(0.37 * lo + 0.11 * hi)
    return I.CSR(A.n_rows, A.n_cols, A.row_ptr, A.cols, A.vals * w)


def _products_equal(p, q, n_rows, vt, seed):
    """p and q give the same bits (x without and, for FP32, with guard zones)."""
    n = p.n_rows_padded
    dt = np.float64 if vt == I.F64 else np.float32
    x = np.concatenate([G.uniform(seed, n_rows), np.zeros(n - n_rows)]).astype(dt)
    ys = []
    for pl in (p, q):
        y = I.DeviceArray(n, dt)
        pl.execute(I.DeviceArray.from_numpy(x), y)
        ys.append(y.numpy())
        if vt == I.F32:
            g = pl.x_guard()
            xg = np.zeros(n + 2 * g, dt)
            xg[g:g + n] = x
            yg = I.DeviceArray(n, dt)
            pl.execute_guarded(I.DeviceArray.from_numpy(xg), g, yg)
            ys.append(yg.numpy())
    for y in ys[1:]:
        np.testing.assert_array_equal(y, ys[0])


@pytest.mark.parametrize("vt,et", TYPES)
@pytest.mark.parametrize("name", ["stencil27_contrast", "levels16", "small27"])
def test_symmetric_refresh(gpu, name, vt, et):
    """New symmetric values for a symmetric plan's pattern, from the host (ispm_plan_update_coefs)
    and on the device (ispm_plan_update_coefs_device), multiply bit for bit like a plan built
    from the new matrix (the planes' repeated rows included: the FP32 guarded kernel reads
    them); host values whose lower entry differs from its mirror are refused."""
    A = _sym_matrix(name)
    A2 = _sym_revalued(A)
    fresh = I.Plan(A2, vt, elm_type=et, flags=I.SPMV_PLAN_SYMMETRIC)
    ph = I.Plan(A, vt, elm_type=et, flags=I.SPMV_PLAN_SYMMETRIC)
    ph.update_coefs(A2)
    _products_equal(ph, fresh, A.n_rows, vt, 51)
    pd = I.Plan(A, vt, elm_type=et, flags=I.SPMV_PLAN_SYMMETRIC)
    pd.update_coefs_device(I.DeviceArray.from_numpy(np.ascontiguousarray(A2.vals, np.float64)))
    _products_equal(pd, fresh, A.n_rows, vt, 52)
    rows = np.repeat(np.arange(A.n_rows), np.diff(A.row_ptr))
    k = int(np.nonzero(A.cols < rows)[0][len(rows) // 3])
    v3 = A2.vals.copy()
    v3[k] *= 1.5
    with pytest.raises(I.IspmError):
        ph.update_coefs(I.CSR(A.n_rows, A.n_cols, A.row_ptr, A.cols, v3))


@pytest.mark.parametrize("name", ["poisson7", "stencil27_contrast", "stencil19_aniso", "small27",
                                  "levels16", "levels131"])
def test_symmetric_guarded_fp32(gpu, name):
    """FP32 products of x held with zero guard zones (ispm_plan_execute_guarded) take the
    four-rows-per-lane kernel: bit-identical to the checked kernel's product, and the CSR
    product within the FP32 bar; a guard below what the plan needs runs the checked kernel."""
    A = _sym_matrix(name)
    p = I.Plan(A, I.F32, elm_type=I.F32, flags=I.SPMV_PLAN_SYMMETRIC)
    g = p.x_guard()
    assert g > 0 and g % 64 == 0
    n = p.n_rows_padded
    x = G.uniform(33, A.n_rows)
    y_plain = run_plan(p, A, x, I.F32)
    for guard in (g, g // 2):
        xg = np.zeros(n + 2 * guard, np.float32)
        xg[guard:guard + A.n_rows] = x
        yd = I.DeviceArray(n, np.float32)
        yd.upload(np.full(n, 7.0, np.float32))
        p.execute_guarded(I.DeviceArray.from_numpy(xg), guard, yd)
        y = yd.numpy()
        np.testing.assert_array_equal(y, y_plain)
        check(A, None, x, y, I.F32, I.F32)


def test_symmetric_plan_integer_exact(gpu):
    A = G.stencil(12, 10, 9, 27, seed=5)
    Ai = I.CSR(A.n_rows, A.n_cols, A.row_ptr, A.cols, np.round(A.vals * 8))
    x = np.round(G.uniform(32, A.n_rows) * 8)
    p = I.Plan(Ai, I.F64, flags=I.SPMV_PLAN_SYMMETRIC)
    check(Ai, None, x, run_plan(p, Ai, x, I.F64), I.F64, I.F64, integer=True)


def test_plan_refuses_mismatched_vectors(mats):
    """The C-ABI takes bare pointers; the Python plan refuses arrays of the wrong element
    type or length before anything is launched."""
    A = mats["poisson7"]
    p = I.Plan(A, I.F64)
    n = p.n_rows_padded
    with pytest.raises(ValueError):
        p.execute(I.DeviceArray(n, np.float32), I.DeviceArray(n, np.float32))
    with pytest.raises(ValueError):
        p.execute(I.DeviceArray(n), I.DeviceArray(n - 1))


def test_plan_default_and_update_coefs(mats):
    A = mats["poisson7"]
    p = I.Plan(A, I.F64)
    inf = p.info()
    assert inf.D == 1 and inf.S == 256
    n = p.n_rows_padded
    x = G.uniform(5, A.n_rows)
    xd = I.DeviceArray.from_numpy(np.concatenate([x, np.zeros(n - A.n_rows)]))
    yd = I.DeviceArray(n)
    p.execute(xd, yd)
    check(A, None, x, yd.numpy(), I.F64, I.F64)
    A2 = I.CSR(A.n_rows, A.n_cols, A.row_ptr, A.cols, A.vals * 3.0 + 1.0)
    p.update_coefs(A2)
    p.execute(xd, yd)
    check(A2, None, x, yd.numpy(), I.F64, I.F64)


@pytest.mark.parametrize("vt,et", [(I.F64, I.F64), (I.F64, I.F32), (I.F32, I.F32)])
def test_full_size_27pt_192(gpu, vt, et):
    """27-point 192^3 contrast matrix at full size (7,077,888 rows): exact A*1 (all values are
    multiples of 1/2, so FP64 sums are exact in any order) and the per-entry FP64 bound on a
    random x against the oracle CSR product."""
    A = _big()
    p = I.Plan(A, vt, elm_type=et)
    inf = p.info()
    # the guess picks 64-row diagonal slices; the shared 27-diagonal pattern moves the
    # default plan to 256-row slices
    assert inf.D == 1 and inf.S == 256 and inf.pattern_planes == 27
    n = p.n_rows_padded
    dt = np.float64 if vt == I.F64 else np.float32
    ones = I.DeviceArray.from_numpy(np.ones(n, dt))
    yd = I.DeviceArray(n, dt)
    p.execute(ones, yd)
    y = yd.numpy()[:A.n_rows].astype(np.float64)
    rs = O.csr_spmv(A, np.ones(A.n_rows))
    if vt == I.F64:
        np.testing.assert_array_equal(y, rs)
    else:
        assert np.all(np.abs(y - rs) <= 1e-5 * O.csr_spmv_abs(A, np.ones(A.n_rows)))
    x = G.uniform(7, A.n_rows)
    xd = I.DeviceArray.from_numpy(np.concatenate([x, np.zeros(n - A.n_rows)]).astype(dt))
    p.execute(xd, yd)
    check(A, None, x, yd.numpy(), vt, et)


@pytest.mark.parametrize("vt,et", [(I.F64, I.F64), (I.F64, I.F32), (I.F32, I.F32)])
def test_full_size_27pt_192_symmetric(gpu, vt, et):
    """The symmetric diagonal plan of the full 27-point 192^3 matrix: exact A*1 in FP64 and
    the per-entry bound on a random x."""
    A = _big()
    p = I.Plan(A, vt, elm_type=et, flags=I.SPMV_PLAN_SYMMETRIC)
    n = p.n_rows_padded
    dt = np.float64 if vt == I.F64 else np.float32
    ones = I.DeviceArray.from_numpy(np.ones(n, dt))
    yd = I.DeviceArray(n, dt)
    p.execute(ones, yd)
    y = yd.numpy()[:A.n_rows].astype(np.float64)
    rs = O.csr_spmv(A, np.ones(A.n_rows))
    if vt == I.F64:
        np.testing.assert_array_equal(y, rs)
    else:
        assert np.all(np.abs(y - rs) <= 1e-5 * O.csr_spmv_abs(A, np.ones(A.n_rows)))
    x = G.uniform(8, A.n_rows)
    xd = I.DeviceArray.from_numpy(np.concatenate([x, np.zeros(n - A.n_rows)]).astype(dt))
    p.execute(xd, yd)
    check(A, None, x, yd.numpy(), vt, et)


_BIG = {}


def _big():
    if "A" not in _BIG:
        _BIG["A"] = G.stencil27_contrast(192)
    return _BIG["A"]


def banded_pairs(nodes=4096, bw=40, per=5, seed=21, far_every=0):
    return G.banded_pairs(nodes, bw, per, seed, far_every)


@pytest.mark.parametrize("vt,et", TYPES)
@pytest.mark.parametrize("S", [64, 256, 512])
@pytest.mark.parametrize("far", [0, 37])
def test_hblock_x_windows(gpu, vt, et, S, far):
    """hblock slices whose columns are close together read x from LDS windows (the plan
    reports how many); the product matches the oracle, with and without far entries that
    leave some slices on the global gathers."""
    A = banded_pairs(far_every=far)
    plan = I.Plan(A, vt, elm_type=et, explicit=dict(S=S, hblock=2))
    inf = plan.info()
    assert inf.hblock == 2 and inf.launch.N == S
    n_slices = inf.device.n_slices
    if far == 0:
        assert inf.window_slices == n_slices
    else:
        # slices holding a far entry keep global gathers where their span is too wide for
        # the vector type's LDS budget or for their own stored slots
        assert inf.window_slices <= n_slices
        if S == 64:
            assert 0 < inf.window_slices < n_slices
    n = plan.n_rows_padded
    dt = np.float64 if vt == I.F64 else np.float32
    x = G.uniform(3, A.n_rows)
    xd = I.DeviceArray.from_numpy(np.concatenate([x, np.zeros(n - A.n_rows)]).astype(dt))
    yd = I.DeviceArray(n, dt)
    plan.execute(xd, yd)
    check(A, None, x, yd.numpy(), vt, et)


def test_hblock_x_windows_integer_exact(gpu):
    A = banded_pairs(nodes=3000, seed=22, far_every=97)
    A.vals = np.round(A.vals * 8)
    plan = I.Plan(A, I.F64, explicit=dict(S=128, hblock=2))
    assert plan.info().window_slices > 0
    n = plan.n_rows_padded
    x = np.round(G.uniform(4, A.n_rows) * 16)
    xd = I.DeviceArray.from_numpy(np.concatenate([x, np.zeros(n - A.n_rows)]))
    yd = I.DeviceArray(n)
    plan.execute(xd, yd)
    np.testing.assert_array_equal(yd.numpy()[:A.n_rows], O.csr_spmv(A, x))


def test_hblock_x_windows_unaligned_x(gpu):
    """x at an address that is not 16-byte aligned: the windows are staged element by
    element; results equal the aligned product bit for bit."""
    A = banded_pairs(nodes=2048, seed=23)
    plan = I.Plan(A, I.F32, elm_type=I.F32, explicit=dict(S=256, hblock=2))
    n = plan.n_rows_padded
    x = G.uniform(5, A.n_rows).astype(np.float32)
    buf = I.DeviceArray.from_numpy(np.concatenate([[0.0], x, np.zeros(n - A.n_rows + 3)]).astype(np.float32))
    xs = I.DeviceArray.from_numpy(np.concatenate([x, np.zeros(n - A.n_rows)]).astype(np.float32))
    y1, y2 = I.DeviceArray(n, np.float32), I.DeviceArray(n, np.float32)
    plan.execute(xs, y1)
    I.check(I.lib().ispm_plan_execute(plan.p, I.C.c_void_p(buf.ptr + 4), y2.vp))
    I.synchronize()
    np.testing.assert_array_equal(y1.numpy(), y2.numpy())
