"""Hand-derived known answers for the SlELL layout rules that the reference's conversion fixes
(formats/conversion.h) and that only invariants checked before: variable slice heights
(estimate_slice_height, conversion.h:6-47, with the min_height raise of 383-384), diagonal
acceptance and rejection (find_diagonals / copy_csr_elts_diags, 234-336) and the format
guess's cut-offs at their edges (csr_to_slell_guess_format, 416-492).  Every expected
slice_ptr, cols and elms entry below is derived by hand in the comments; nothing here reruns
the conversion's formulas."""
import numpy as np
import pytest

import ispm_amd as I


def csr_from_rows(rows, n_cols, val=lambda r, c: 100.0 * r + c + 1):
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    cols = np.concatenate([np.asarray(r, np.int32) for r in rows]) if rp[-1] else np.zeros(0, np.int32)
    vals = np.array([val(r, c) for r, row in enumerate(rows) for c in row], np.float64)
    return I.CSR(len(rows), n_cols, rp, cols, vals)


def records(hs, P):
    return hs.slice_ptr[:(hs.n_slices + 1) * P].reshape(hs.n_slices + 1, P).astype(np.int64)


# --------------------------------------------------------------- variable heights ---------

def test_var_height_waste_stop():
    """512 rows of one entry except row 40 with three; cap (requested height) 64.
    waste_per_row = 0.01 * nnz * 100 / n = 514 / 512 = 1.0039.
    Slice at row 0: rows 0..39 add no waste; row 40 (slice_size 40) would add 40 * (3 - 1) =
    80 > 1.0039 * 40 -> stop at 40, rounded down to 32.
    Slice at row 32: row 40 (slice_size 8, below 32: only the element cap 4 * 64 = 256
    applies, 11 + 16 <= 256) widens it to 3, waste 16; every later row adds 2, and at slice
    size 32 the waste 16 + 2 * 24 = 64 > 1.0039 * 32 -> stop at 32 (a power of two).
    Then only single-entry rows: slices reach the cap, 64 rows each, to row 512.
    Elements: 32 * 1, then 32 * 3, then 64 * 1 per slice."""
    rows = [[r] for r in range(512)]
    rows[40] = [0, 1, 2]
    A = csr_from_rows(rows, 512)
    hs = I.csr_to_slell(A, 64, var_height=True)
    rec = records(hs, 2)
    assert list(rec[:, 1]) == [0, 32, 64, 128, 192, 256, 320, 384, 448, 512]
    assert list(rec[:, 0]) == [0, 32, 128, 192, 256, 320, 384, 448, 512, 576]
    # row 40 is row 8 of slice 1 (elements at 32 + 8 + k * 32); its neighbours pad with the
    # column of the row above (row 41 repeats row 40's columns, row 33 those of row 32)
    assert [hs.cols[32 + 8 + 32 * k] for k in range(3)] == [0, 1, 2]
    assert [hs.cols[32 + 9 + 32 * k] for k in range(3)] == [41, 1, 2]
    assert [hs.cols[32 + 1 + 32 * k] for k in range(3)] == [33, 0, 0]
    assert [hs.elms[32 + 9 + 32 * k] for k in range(3)] == [4142.0, 0.0, 0.0]


def test_var_height_element_cap_round_down_and_min_height():
    """Rows 0..63 hold 9 entries, row 64 holds 300, rows 65..511 one; cap 64 (4 * 64 = 256
    elements), min_height 4.
    Rows of 9: the slice stops when 9 * (k + 1) > 256, at k = 28 -> 16 rows; four such slices
    cover rows 0..63 (the fourth stops at row 64: 144 + 300 + 16 * 291 > 256).
    Row 64 alone is 300 > 256 elements: slice_size 1, rounded 1, raised to min_height 4:
    rows 64..67, 300 wide.
    Then single-entry rows fill 64-row slices until the rows run out: at row 452, 60 rows
    remain -> 32; at 484, 28 -> 16; at 500, 12 -> 8; at 508, 4 -> 4."""
    rows = [list(range(r % 400, r % 400 + 9)) for r in range(64)]
    rows.append(list(range(0, 300)))
    rows += [[r] for r in range(65, 512)]
    A = csr_from_rows(rows, 512)
    hs = I.csr_to_slell(A, 64, var_height=True, min_height=4)
    rec = records(hs, 2)
    assert list(rec[:, 1]) == [0, 16, 32, 48, 64, 68, 132, 196, 260, 324, 388, 452, 484, 500, 508, 512]
    assert list(rec[:, 0]) == [0, 144, 288, 432, 576, 1776, 1840, 1904, 1968, 2032, 2096, 2160,
                               2192, 2208, 2216, 2220]
    # slice 4 (rows 64..67, height 4): row 65's 299 padding slots take row 64's columns
    base = 576
    assert hs.cols[base + 1] == 65 and hs.cols[base + 1 + 4 * 7] == 7
    assert hs.elms[base + 1 + 4 * 7] == 0.0
    assert hs.cols[base + 0 + 4 * 299] == 299 and hs.elms[base + 4 * 299] == 6400.0 + 299 + 1


# ------------------------------------------------------------------ diagonals -------------

def tridiagonal_rest(rows, n):
    for r in range(len(rows), n):
        rows.append([c for c in (r - 1, r, r + 1) if 0 <= c < n])
    return rows


def test_diagonal_accepts_off_stencil_entry_of_the_only_longest_row():
    """Slice 0 (height 4) of a tridiagonal matrix whose row 2 also holds column 7:
    row 0 {0,1}, row 1 {0,1,2}, row 2 {1,2,3,7} (longest), row 3 {2,3,4}.
    Row 2's offsets -1, 0, +1 are shared by every row of the current maximal length and are
    accepted (maxlen 4 -> 1); offset +5 is then missing only from rows whose remaining length
    (0) is below that maximum, so it is accepted too: diagonal positions (col - 2 + 4) 3, 4,
    5, 9 and no ELL columns.  Planes (4 rows each): -1: [0, a10, a21, a32], 0: [a00..a33],
    +1: [a01, a12, a23, a34], +5: [0, 0, a27, 0]."""
    n = 512
    rows = tridiagonal_rest([[0, 1], [0, 1, 2], [1, 2, 3, 7], [2, 3, 4]], n)
    A = csr_from_rows(rows, n)
    hs = I.csr_to_slell(A, 4, diags=True)
    rec = records(hs, 3)
    assert list(rec[0]) == [0, 0, (hs.n_slices + 1) * 3]
    assert list(rec[1, :2]) == [16, 0]
    d0, d1 = rec[0, 2], rec[1, 2]
    assert list(hs.slice_ptr[d0:d1]) == [3, 4, 5, 9]
    a = lambda r, c: 100.0 * r + c + 1
    want = [0, a(1, 0), a(2, 1), a(3, 2),
            a(0, 0), a(1, 1), a(2, 2), a(3, 3),
            a(0, 1), a(1, 2), a(2, 3), a(3, 4),
            0, 0, a(2, 7), 0]
    assert list(hs.elms[:16]) == want
    # every later slice is plain tridiagonal, no ELL part; positions are column - local row
    # + S, so slice 1 (rows 4..7, row 4 = {3,4,5}) has 7, 8, 9
    assert list(hs.slice_ptr[rec[1, 2]:rec[2, 2]]) == [7, 8, 9]
    assert rec[2, 1] == 0


def test_diagonal_rejected_when_an_equally_long_row_lacks_it():
    """As above but row 3 = {2,3,4,5} is as long as row 2 (the first longest row is taken).
    After -1, 0, +1 are accepted both rows have one entry left (the maximum, 1); row 3 lacks
    offset +5, so it is rejected: diagonal positions 3, 4, 5 and one ELL plane holding row 2's
    column 7 and row 3's column 5; rows 0 and 1 pad with the column of the row above (row 0:
    column 0).  elms: the ELL plane [0, 0, a27, a35] first, then the three diagonal planes."""
    n = 512
    rows = tridiagonal_rest([[0, 1], [0, 1, 2], [1, 2, 3, 7], [2, 3, 4, 5]], n)
    A = csr_from_rows(rows, n)
    hs = I.csr_to_slell(A, 4, diags=True)
    rec = records(hs, 3)
    assert list(rec[1, :2]) == [16, 4]
    assert list(hs.slice_ptr[rec[0, 2]:rec[1, 2]]) == [3, 4, 5]
    assert list(hs.cols[:4]) == [0, 0, 7, 5]
    a = lambda r, c: 100.0 * r + c + 1
    want = [0, 0, a(2, 7), a(3, 5),
            0, a(1, 0), a(2, 1), a(3, 2),
            a(0, 0), a(1, 1), a(2, 2), a(3, 3),
            a(0, 1), a(1, 2), a(2, 3), a(3, 4)]
    assert list(hs.elms[:16]) == want


# ------------------------------------------------------------------- the guess ------------
# 3200 rows: the guess examines (3200 + 99) / 100 = 32 rows (one 32-row trial slice) and
# counts the elements of rows 0..32 -- row_ptr[rows_to_examine + 1], one row more than the
# trial slice (conversion.h:425).

N = 3200


def rest(rows):
    return rows + [[r] for r in range(len(rows), N)]


@pytest.mark.parametrize("len32, var", [(28, False), (27, True)])
def test_guess_waste_ratio_edge(len32, var):
    """Row 0 holds 3 entries, rows 1..31 one: waste 31 * 2 = 62.  With 28 entries in row 32
    the examined elements are 3 + 31 + 28 = 62 and the waste ratio is exactly 100 (not above
    100: fixed height); with 27 they are 61, ratio 101: variable height (256 rows)."""
    rows = [[0, 1, 2]] + [[r] for r in range(1, 32)] + [list(range(100, 100 + len32))]
    p = I.csr_to_slell_guess(csr_from_rows(rest(rows), N)).params()
    assert p["V"] is var
    # fixed: 2048 / (62 / 32) = 1057 -> 1024 -> clamped to 512
    assert p["S"] == (256 if var else 512)


@pytest.mark.parametrize("len32, S", [(0, 256), (1, 128)])
def test_guess_slice_height_edge(len32, S):
    """Rows 0..31 hold 8 entries: 256 (+ row 32) examined elements over 32 rows.  256 -> 8
    per row, 2048 / 8 = 256 exactly -> 256; 257 -> 2048 / 8.03 = 255 -> rounded down to 128."""
    rows = [list(range(40 * r % 3000, 40 * r % 3000 + 15, 2)) for r in range(32)]
    rows.append(list(range(3100, 3100 + len32)))
    p = I.csr_to_slell_guess(csr_from_rows(rest(rows), N)).params()
    assert p["S"] == S and p["V"] is False


@pytest.mark.parametrize("width, S", [(100, 32), (0, 512)])
def test_guess_slice_height_clamps(width, S):
    """100 entries per row: 2048 / 100 = 20 -> 16, raised to the minimum 32.  No entries in
    the examined rows (1 element counted at least): 65536, lowered to the maximum 512."""
    rows = [list(range(r % 2000, r % 2000 + width)) for r in range(33)]
    p = I.csr_to_slell_guess(csr_from_rows(rest(rows), N)).params()
    assert p["S"] == S


@pytest.mark.parametrize("short, diags", [(False, False), (True, True)])
def test_guess_diagonal_share_edge(short, diags):
    """Every trial row holds its diagonal and one entry of its own (column r + 1000 + 3r):
    find_diagonals accepts the main diagonal only (1 position -> 32 diagonal elements).
    Examined elements 64 -> 32 > 64 * 50 / 100 = 32 is false: no diagonals; with row 5
    holding only its diagonal, 63 -> 32 > 31: diagonal slices."""
    rows = [[r, 4 * r + 1000] for r in range(32)]
    if short:
        rows[5] = [5]
    rows.append([])
    p = I.csr_to_slell_guess(csr_from_rows(rest(rows), N)).params()
    assert p["D"] is diags and p["hblock"] == 1


@pytest.mark.parametrize("len32, hblock", [(64, 1), (63, 2)])
def test_guess_pair_share_edge(len32, hblock):
    """Every trial row is one aligned column pair (2k, 2k+1) at its own position, so no
    diagonal is shared: 32 pairs.  Examined elements 64 + row 32: with 64 entries in row 32,
    32 > 128 * 25 / 100 = 32 is false (hblock 1); with 63, 32 > 31: hblock 2."""
    rows = [[2 * (37 * r % 1500), 2 * (37 * r % 1500) + 1] for r in range(32)]
    rows.append(list(range(3100 - len32, 3100)))
    p = I.csr_to_slell_guess(csr_from_rows(rest(rows), N)).params()
    assert p["hblock"] == hblock and p["D"] is False


@pytest.mark.parametrize("bad", ["unsorted", "column_range", "row_ptr"])
def test_conversion_rejects_malformed_csr(bad):
    """CSR rows must be column-sorted (the diagonal encoder would otherwise place an entry
    past its slice's planes), columns in range and row_ptr non-decreasing: every conversion
    refuses such input with ISPM_ERR_INVALID instead of writing out of bounds."""
    rows = tridiagonal_rest([], 600)
    A = csr_from_rows(rows, 600)
    if bad == "unsorted":
        A.cols[2], A.cols[3] = A.cols[3], A.cols[2]          # row 1: 1, 0, 2
    elif bad == "column_range":
        A.cols[7] = 600
    else:
        A.row_ptr[10] = A.row_ptr[9] - 1
    for call in (lambda: I.csr_to_slell(A, 4, diags=True), lambda: I.csr_to_slell(A, 64),
                 lambda: I.csr_to_slell_guess(A),
                 lambda: I.Plan(A, I.F64, flags=I.SPMV_PLAN_NO_DEVICE)):
        with pytest.raises(I.IspmError) as e:
            call()
        assert e.value.code == I.ERR_INVALID
