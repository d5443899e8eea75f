"""oracle.py -- ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the benchmark's
cpu_baseline leg as the checker; the product library never uses it.

What pins each oracle function (tests/test_oracle.py):
  csr_spmv          brute-force dense products (numpy) on small matrices; exact equality on
                    integer-valued inputs; stencil closed form A*1 = row sums.
  slell_decode      dense(decode(csr_to_slell(A))) == A for every variant; hand-built layouts.
  dot / ppcg_*      closed forms on hand cases (n = 4); the pipelined recurrences are checked
                    against a plain-CG run on the same system (same iterates in exact
                    arithmetic) and against numpy's dense solve.
  pcg / ppcg        dense solve (numpy.linalg.solve) on small SPD systems, residual tests.
  fastainv_sym      closed forms: diagonal A -> D = 1/a_ii, W = I; 2x2 exact inverse;
                    droptol = 0 without truncation -> Ainv2*Ainv1 = A^-1 (dense check);
                    invariants (unit-triangular W, symmetric M, sorted rows, reversal
                    structure of Ainv1/Ainv2); an independent dense right-looking AINV with
                    the same dropping rule for droptol > 0; storage truncation by max_elts.
  droptol tuner     the reference's own DroptolTuner compiled from its header
                    (oracle/_ref/libref_droptol.so) on identical time sequences, and the
                    closed-form golden-section probe sequence.
  ldu_dense         lduMatrix (OpenFOAM owner/neighbour face addressing) to a dense matrix,
                    entry by entry in pure Python: a hand-built 3-cell case.
  foam_residual     OpenFOAM's normalised residual (lduMatrix::solver::normFactor): a hand-
                    computed 2x2 case; invariance under scaling A and b together; for a
                    constant x the first term of the norm factor vanishes (A x = sumA xbar),
                    closed form; zero at the exact solution.  The foam-extend caller is not
                    available, so its use of this measure is parity unpinned; the formula is
                    OpenFOAM's.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} not built: run `make -C {_HERE}`")
        L = C.CDLL(path)
        L.orc_set_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_csr_spmv.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp]
        L.orc_csr_spmv_abs.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp]
        L.orc_slell_decode.restype = C.c_longlong
        L.orc_slell_decode.argtypes = [C.c_int, C.c_uint, C.c_uint, C.c_uint, C.c_int, C.c_int,
                                       C.POINTER(C.c_uint), C.POINTER(C.c_uint), _dp, C.c_longlong,
                                       _ip, C.POINTER(C.c_longlong), _dp]
        L.orc_dot.restype = C.c_double
        L.orc_dot.argtypes = [_dp, _dp, C.c_int]
        L.orc_ppcg_update_scalars.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double]
        L.orc_ppcg_update_vectors.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double] + [_dp] * 10 + [C.c_int]
        L.orc_copy_indexed.argtypes = [_dp, _dp, _ip, C.c_int]
        L.orc_add_indexed.argtypes = [_dp, _dp, _ip, C.c_int]
        for f in (L.orc_pcg, L.orc_ppcg):
            f.restype = C.c_int
        L.orc_pcg.argtypes = [C.c_int, _ip, _ip, _dp, _ip, _ip, _dp, _ip, _ip, _dp, _dp, _dp,
                              C.c_double, C.c_int, _dp]
        L.orc_ppcg.argtypes = [C.c_int, _ip, _ip, _dp, _ip, _ip, _dp, _ip, _ip, _dp, _dp, _dp,
                               C.c_double, C.c_int, _dp, _dp]
        L.orc_fastainv_sym.restype = C.c_void_p
        L.orc_fastainv_sym.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, C.c_longlong, C.c_double]
        L.orc_fastainv_nnz.restype = C.c_int
        L.orc_fastainv_nnz.argtypes = [C.c_void_p, C.c_int]
        L.orc_fastainv_copy.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _dp]
        L.orc_fastainv_free.argtypes = [C.c_void_p]
        for name, res, args in (("new", C.c_void_p, [C.c_double, C.c_double]),
                                ("get", C.c_double, [C.c_void_p]),
                                ("best", C.c_double, [C.c_void_p]),
                                ("note", None, [C.c_void_p, C.c_double]),
                                ("restart", None, [C.c_void_p]),
                                ("reset_phase", None, [C.c_void_p]),
                                ("iters", C.c_int, [C.c_void_p]),
                                ("free", None, [C.c_void_p])):
            f = getattr(L, "orc_dtt_" + name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref_droptol_lib():
    """The reference's own DroptolTuner (oracle/_ref), or None if it was not built."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libref_droptol.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        for name, res, args in (("new", C.c_void_p, [C.c_double, C.c_double]),
                                ("get", C.c_double, [C.c_void_p]),
                                ("best", C.c_double, [C.c_void_p]),
                                ("note", None, [C.c_void_p, C.c_double]),
                                ("restart", None, [C.c_void_p]),
                                ("reset_phase", None, [C.c_void_p]),
                                ("iters", C.c_int, [C.c_void_p]),
                                ("free", None, [C.c_void_p])):
            f = getattr(L, "ref_dtt_" + name)
            f.restype = res
            f.argtypes = args
        _ref = L
    return _ref


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def P(a, t=_dp):
    return a.ctypes.data_as(t)


def set_threads(n=0):
    return lib().orc_set_threads(int(n))


def csr_spmv(A, x):
    rp, ci, v, x = _i(A.row_ptr), _i(A.cols), _d(A.vals), _d(x)
    y = np.empty(A.n_rows)
    lib().orc_csr_spmv(A.n_rows, P(rp, _ip), P(ci, _ip), P(v), P(x), P(y))
    return y


def csr_spmv_abs(A, x):
    rp, ci, v, x = _i(A.row_ptr), _i(A.cols), _d(A.vals), _d(x)
    y = np.empty(A.n_rows)
    lib().orc_csr_spmv_abs(A.n_rows, P(rp, _ip), P(ci, _ip), P(v), P(x), P(y))
    return y


def slell_decode(hs):
    """(rows, cols, vals) of every stored slot of a HostSlell."""
    sp = np.ascontiguousarray(hs.slice_ptr, dtype=np.uint32)
    cl = np.ascontiguousarray(hs.cols, dtype=np.uint32)
    el = _d(hs.elms_f64())
    cap = int(el.size) + 16
    r = np.empty(cap, np.int32)
    c = np.empty(cap, np.int64)
    v = np.empty(cap)
    up = C.POINTER(C.c_uint)
    n = lib().orc_slell_decode(hs.n_cols, hs.n_slices, hs.slice_height, hs.hblock, int(hs.var_height),
                               int(hs.diags), P(sp, up), P(cl, up), P(el), cap, P(r, _ip),
                               P(c, C.POINTER(C.c_longlong)), P(v))
    if n < 0:
        raise ValueError(f"slell_decode failed ({n})")
    return r[:n], c[:n], v[:n]


def slell_spmv(hs, x):
    """y = A x computed from the stored slots of a SlELL layout, x read as 0 outside
    [0, n_cols) (the texture semantics of cuda/texture-wrap.h)."""
    r, c, v = slell_decode(hs)
    x = _d(x)
    ok = (c >= 0) & (c < hs.n_cols)
    y = np.zeros(hs.n_rows)
    np.add.at(y, r[ok], v[ok] * x[c[ok]])
    return y


def dot(x, y):
    x, y = _d(x), _d(y)
    return lib().orc_dot(P(x), P(y), x.size)


def ppcg_update_scalars(alpha, beta_unused, gamma, gammaold, delta):
    a = C.c_double(alpha)
    b = C.c_double(0)
    lib().orc_ppcg_update_scalars(C.byref(a), C.byref(b), gamma, gammaold, delta)
    return a.value, b.value


def ppcg_update_vectors(alpha, beta, n, m, p, s, q, z, x, r, u, w):
    """In-place on copies; returns (resnorm, gamma, delta, p, s, q, z, x, r, u, w)."""
    vs = [_d(v).copy() for v in (p, s, q, z, x, r, u, w)]
    n_, m_ = _d(n), _d(m)
    rn, g, d = C.c_double(), C.c_double(), C.c_double()
    lib().orc_ppcg_update_vectors(C.byref(rn), C.byref(g), C.byref(d), alpha, beta, P(n_), P(m_),
                                  *[P(v) for v in vs], n_.size)
    return (rn.value, g.value, d.value, *vs)


def _csr_args(M):
    if M is None:
        return None, None, None, ()
    rp, ci, v = _i(M.row_ptr), _i(M.cols), _d(M.vals)
    return P(rp, _ip), P(ci, _ip), P(v), (rp, ci, v)


def pcg(A, b, M1=None, M2=None, tol=1e-8, maxit=10000, x0=None, pipelined=False):
    """(x, iterations, true relative residual[, L1 resnorm]); iterations = -1 if maxit hit."""
    a = _csr_args(A)
    m1 = _csr_args(M1)
    m2 = _csr_args(M2)
    b = _d(b)
    x = np.zeros(A.n_rows) if x0 is None else _d(x0).copy()
    rel = C.c_double()
    if pipelined:
        l1 = C.c_double()
        it = lib().orc_ppcg(A.n_rows, a[0], a[1], a[2], m1[0], m1[1], m1[2], m2[0], m2[1], m2[2],
                            P(b), P(x), tol, maxit, C.byref(rel), C.byref(l1))
        return x, it, rel.value, l1.value
    it = lib().orc_pcg(A.n_rows, a[0], a[1], a[2], m1[0], m1[1], m1[2], m2[0], m2[1], m2[2],
                       P(b), P(x), tol, maxit, C.byref(rel))
    return x, it, rel.value


def fastainv_sym(A, out_double=False, max_elts=8192000, droptol=0.125):
    """Restated fastainv_sym: returns ((rp1, ci1, v1), (rp2, ci2, v2)) or None for n <= 0."""
    rp, ci, v = _i(A.row_ptr), _i(A.cols), _d(A.vals)
    h = lib().orc_fastainv_sym(int(out_double), A.n_rows, P(rp, _ip), P(ci, _ip), P(v),
                               int(max_elts), float(droptol))
    if not h:
        return None
    out = []
    try:
        for which in (1, 2):
            nnz = lib().orc_fastainv_nnz(h, which)
            r = np.empty(A.n_rows + 1, np.int32)
            c = np.empty(nnz, np.int32)
            x = np.empty(nnz)
            lib().orc_fastainv_copy(h, which, P(r, _ip), P(c, _ip), P(x))
            out.append((r, c, x))
    finally:
        lib().orc_fastainv_free(h)
    return tuple(out)


class DroptolTuner:
    """Restated DroptolTuner (or, with ref=True, the reference's own compiled class)."""

    def __init__(self, lo=-10.0, hi=-3.0, ref=False):
        self.L = ref_droptol_lib() if ref else lib()
        if self.L is None:
            raise RuntimeError("reference DroptolTuner not built (oracle/_ref)")
        self.pre = "ref_dtt_" if ref else "orc_dtt_"
        self.h = self._f("new")(lo, hi)

    def _f(self, n):
        return getattr(self.L, self.pre + n)

    def get(self):
        return self._f("get")(self.h)

    def best(self):
        return self._f("best")(self.h)

    def note(self, t):
        self._f("note")(self.h, float(t))

    def restart(self):
        self._f("restart")(self.h)

    def reset_phase(self):
        self._f("reset_phase")(self.h)

    def iterations(self):
        return self._f("iters")(self.h)

    def __del__(self):
        try:
            self._f("free")(self.h)
        except Exception:
            pass


def ldu_dense(n, lower_addr, upper_addr, diag, upper, lower=None):
    """Dense matrix of an lduMatrix: A[i][i] = diag[i]; face f: A[l][u] = upper[f],
    A[u][l] = lower[f] (upper when lower is None)."""
    D = np.zeros((n, n))
    for i in range(n):
        D[i, i] = diag[i]
    for f in range(len(lower_addr)):
        l, u = int(lower_addr[f]), int(upper_addr[f])
        D[l, u] = upper[f]
        D[u, l] = (upper if lower is None else lower)[f]
    return D


def foam_residual(A, b, x0, x):
    """(normFactor, residual): OpenFOAM's normalised residual of x, with the norm factor of
    the initial x0: xbar = mean(x0), pA = (A 1) xbar,
    normFactor = sum |A x0 - pA| + |b - pA| + 1e-20, residual = sum |b - A x| / normFactor."""
    n = A.n_rows
    sumA = csr_spmv(A, np.ones(n))
    pA = sumA * (np.sum(x0) / n)
    nf = float(np.sum(np.abs(csr_spmv(A, x0) - pA) + np.abs(b - pA))) + 1e-20
    return nf, float(np.sum(np.abs(b - csr_spmv(A, x)))) / nf

