"""Host CSR -> SlELL conversion: layout invariants and the reference's layout rules
(formats/conversion.h), checked on the CPU."""
import numpy as np
import pytest
import scipy.sparse as sp

import ispm_amd as I
import ispm_amd.gen as G
import oracle as O


def P_of(hs):
    return 1 + (hs.hblock > 1) + int(hs.var_height) + 2 * int(hs.diags)


def slice_bounds(hs):
    P = P_of(hs)
    rec = hs.slice_ptr[: (hs.n_slices + 1) * P].reshape(hs.n_slices + 1, P).astype(np.int64)
    if hs.var_height:
        y = rec[:, 1]
    else:
        y = np.arange(hs.n_slices + 1) * hs.slice_height
    return rec, y


def test_hblock_cost_quirk_known_answer():
    """Slice with per-row (singles, blocks) = (0,4), (8,1), (4,0), (2,1).  With the reference's
    argument order in the ternary search (conversion.h:152-154) the search keeps
    minblen = 2 and the final scan picks 2 block planes (7 single planes): 4*(7+2) = 36
    column slots and 4*7 + 2*2*4 = 44 element slots.  A search with both costs evaluated
    alike would pick 1 block plane (8 single planes, 40 element slots)."""
    rows = [[0, 1, 2, 3, 4, 5, 6, 7],
            [1, 3, 5, 7, 9, 11, 13, 15, 20, 21],
            [1, 3, 5, 7],
            [1, 3, 8, 9]]
    rp = np.cumsum([0] + [len(r) for r in rows])
    A = I.CSR(4, 22, rp, np.concatenate(rows), np.arange(1, rp[-1] + 1, dtype=np.float64))
    hs = I.csr_to_slell(A, 4, hblock=2)
    assert hs.slice_ptr[0] == 0 and hs.slice_ptr[1] == 0
    assert hs.slice_ptr[2] == 44 and hs.slice_ptr[3] == 36
    r, c, v = O.slell_decode(hs)
    ok = v != 0
    got = sp.csr_matrix((v[ok], (r[ok], c[ok])), shape=(hs.n_rows, 22))
    assert abs(got[:4] - A.to_scipy()).max() == 0


def test_ell_padding_takes_column_of_row_above():
    A = G.random_sparse(700, 3, min_len=0, max_len=12)
    hs = I.csr_to_slell(A, 32)
    rec, y = slice_bounds(hs)
    rp = A.row_ptr
    for s in range(hs.n_slices):
        e0, e1 = rec[s, 0], rec[s + 1, 0]
        S = 32
        w = (e1 - e0) // S
        for r in range(S):
            row = s * S + r
            L = rp[row + 1] - rp[row] if row < A.n_rows else 0
            for k in range(L, w):
                slot = e0 + r + k * S
                want = hs.cols[slot - 1] if r > 0 else 0
                assert hs.cols[slot] == want and hs.elms[slot] == 0


def test_padding_to_512_rows_and_input_untouched():
    A = G.random_sparse(1000, 4, max_len=9)
    rp0, c0, v0 = A.row_ptr.copy(), A.cols.copy(), A.vals.copy()
    hs = I.csr_to_slell(A, 64)
    assert hs.n_rows == 1024 and hs.n_cols == 1000 and hs.n_slices == 16
    np.testing.assert_array_equal(A.row_ptr, rp0)
    np.testing.assert_array_equal(A.cols, c0)
    np.testing.assert_array_equal(A.vals, v0)


@pytest.mark.parametrize("fmt", [dict(S=64), dict(S=32, D=True), dict(S=128, V=True),
                                 dict(S=256, V=True, D=True), dict(S=64, hblock=2),
                                 dict(S=64, hblock=2, D=True)])
def test_slice_records(fmt):
    A = G.random_sparse(2000, 5, min_len=0, max_len=80)
    hs = I.csr_to_slell(A, fmt["S"], fmt.get("hblock", 1), fmt.get("V", False), fmt.get("D", False))
    if fmt.get("hblock", 1) > 1:
        assert not hs.diags            # hblock > 1 disables diagonal slices
    P = P_of(hs)
    rec, y = slice_bounds(hs)
    assert np.all(np.diff(rec[:, 0]) >= 0)
    assert y[0] == 0 and y[-1] == hs.n_rows
    h = np.diff(y)
    assert np.all(h >= 4) and np.all((h & (h - 1)) == 0)
    if not hs.var_height:
        assert np.all(h == fmt["S"])
    else:
        assert np.all(h <= fmt["S"])
    if hs.diags:
        d = rec[:, 2 + int(hs.var_height)]
        assert d[0] == (hs.n_slices + 1) * P and d[-1] == hs.slice_ptr.size
    assert rec[-1, 0] == hs.elms.size


def test_guess_decisions():
    A = G.poisson7(32)
    assert I.csr_to_slell_guess(A).params() == dict(S=256, hblock=1, V=False, D=True)
    B = G.stencil27_contrast(32)
    assert I.csr_to_slell_guess(B).params() == dict(S=64, hblock=1, V=False, D=True)
    # alternating very short and very long rows: padding waste > 100% -> variable height
    n = 4096
    lens = np.where(np.arange(n) % 8 == 0, 200, 2)
    rp = np.concatenate([[0], np.cumsum(lens)])
    cols = np.concatenate([np.sort(np.random.default_rng(i).choice(n, L, replace=False)) for i, L in enumerate(lens)])
    C = I.CSR(n, n, rp, cols, np.ones(rp[-1]))
    assert I.csr_to_slell_guess(C).params()["V"] is True
    # rows made of aligned column pairs at random positions: no diagonals -> hblock 2
    rows = []
    rng = np.random.default_rng(1)
    for r in range(n):
        base = np.sort(rng.choice(n // 2, 6, replace=False)) * 2
        rows.append(np.sort(np.concatenate([base, base + 1])))
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])])
    Dm = I.CSR(n, n, rp, np.concatenate(rows), np.ones(rp[-1]))
    p = I.csr_to_slell_guess(Dm).params()
    assert p["hblock"] == 2 and p["D"] is False
    # 2048 / 12 elements per row -> 170 -> power of two below: 128
    assert p["S"] == 128


def test_diagonal_positions_of_stencil():
    A = G.poisson7(16)
    hs = I.csr_to_slell(A, 64, diags=True)
    rec, y = slice_bounds(hs)
    # an interior slice (rows 2048..2111: z=8, y=0..3) has the 7 stencil diagonals except
    # none missing -> diag_pos - S = offsets {-256,-16,-1,0,1,16,256} relative to its first row
    s = 2048 // 64 + 1
    d0, d1 = rec[s, 2], rec[s + 1, 2]
    offs = hs.slice_ptr[d0:d1].astype(np.int64) - 64 - s * 64
    assert list(offs) == [-256, -16, -1, 0, 1, 16, 256]


def test_bf16_and_f32_narrowing():
    A = G.random_sparse(500, 6, max_len=10)
    h32 = I.csr_to_slell(A, 64, out_type=I.F32)
    hbf = I.csr_to_slell(A, 64, out_type=I.BF16)
    assert h32.elms.dtype == np.float32 and hbf.elms.dtype == np.uint16
    r, c, v = O.slell_decode(hbf)
    ok = v != 0
    got = sp.csr_matrix((v[ok], (r[ok], c[ok])), shape=(hbf.n_rows, hbf.n_cols))[:A.n_rows].toarray()
    ref = A.to_scipy().toarray()
    m = ref != 0
    assert np.all(np.abs(got[m] - ref[m]) <= 2.0 ** -8 * np.abs(ref[m]))


def test_fixed_height_below_min_height_rejected():
    """A fixed slice height below min_height would give slices of min_height rows that the
    recorded height does not describe (conversion.h:383-384): rejected."""
    A = G.random_sparse(300, 31, min_len=0, max_len=5)
    with pytest.raises(I.IspmError):
        I.csr_to_slell(A, 2, min_height=4)
    for S in (1, 2):
        hs = I.csr_to_slell(A, S)          # min_height follows S
        assert hs.slice_height == S and hs.n_slices * S == hs.n_rows
