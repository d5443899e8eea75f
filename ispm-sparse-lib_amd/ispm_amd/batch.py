"""Multi-RHS block PCG and the persistent batched small-system solver (include/ispm.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import CSR, F64, PLAN_MIN_BYTES, SPMV_PLAN_SYMMETRIC, DeviceArray, Plan, check, lib
from .solver import PCG, PREC_FP64, SolveStats, SolverOpts


def block_solve(A: CSR, Bm, M=None, tol=1e-10, maxit=100000, check_every=16, symmetric=False,
                use_graph=True):
    """Solve A X[j] = Bm[j] for the k rows of Bm (k <= 8) as one block of PCG recurrences.
    symmetric: A as a symmetric diagonal plan (ISPM_SPMV_PLAN_SYMMETRIC; the factors' plans
    must then pad to the same row count, as for n a multiple of 512).  use_graph: replay
    batches of check_every iterations from a captured hipGraph (else eager launches).
    Returns (X as a k x n array, list of per-column stats dicts)."""
    Bm = np.atleast_2d(np.asarray(Bm, np.float64))
    k = Bm.shape[0]
    Ap = Plan(A, F64, elm_type=F64, flags=SPMV_PLAN_SYMMETRIC if symmetric else 0)
    Mp = None
    if M is not None:
        Mp = (Plan(M[0], F64, elm_type=F64, kind=PLAN_MIN_BYTES), Plan(M[1], F64, elm_type=F64, kind=PLAN_MIN_BYTES))
    n = Ap.n_rows_padded
    opts = SolverOpts(PCG, PREC_FP64, tol, maxit, check_every, int(use_graph), 0.1)
    s = C.c_void_p()
    check(lib().ispm_block_solver_create(C.byref(s), Ap.p, Mp[0].p if Mp else None, Mp[1].p if Mp else None,
                                         k, C.byref(opts)))
    try:
        Bp = np.zeros((k, n))
        Bp[:, :A.n_rows] = Bm
        bd = DeviceArray.from_numpy(Bp.ravel())
        xd = DeviceArray(k * n)
        stats = (SolveStats * k)()
        code = lib().ispm_block_solver_solve(s, bd.vp, n, xd.vp, n, stats)
        if code not in (0, 6):
            check(code)
        X = xd.numpy().reshape(k, n)[:, :A.n_rows]
        out = [dict(iterations=st.iterations, replacements=st.replacements, converged=bool(st.converged),
                    relres=st.relres) for st in stats]
        return X, out
    finally:
        lib().ispm_block_solver_destroy(s)


class Batch:
    """A set of independent FP32 SPD systems uploaded for the persistent batched solver."""

    def __init__(self, systems):
        rows = np.array([A.n_rows for A in systems], np.int32)
        offs = np.concatenate([[0], np.cumsum(rows)])
        rp, cols, vals = [np.zeros(1, np.int64)], [], []
        base = 0
        for A, o in zip(systems, offs[:-1]):
            rp.append(A.row_ptr[1:].astype(np.int64) + base)
            cols.append(A.cols.astype(np.int64) + o)
            vals.append(A.vals.astype(np.float32))
            base += A.nnz
        self.row_ptr = np.concatenate(rp).astype(np.int32)
        self.cols = np.concatenate(cols).astype(np.int32)
        self.vals = np.concatenate(vals).astype(np.float32)
        self.rows = rows
        b = C.c_void_p()
        check(lib().ispm_batch_create(C.byref(b), len(systems), rows.ctypes.data, self.row_ptr.ctypes.data,
                                      self.cols.ctypes.data, self.vals.ctypes.data))
        self.b = b
        total = C.c_longlong()
        self.row_off = np.zeros(len(systems), np.int64)
        check(lib().ispm_batch_layout(b, C.byref(total), self.row_off.ctypes.data))
        self.rows_padded = total.value
        mb = C.c_longlong()
        check(lib().ispm_batch_matrix_bytes(b, C.byref(mb)))
        self.matrix_bytes = mb.value   # stored values, columns, slice offsets and widths

    def pack(self, vecs):
        out = np.zeros(self.rows_padded, np.float32)
        for v, o in zip(vecs, self.row_off):
            out[o:o + v.size] = v
        return out

    def unpack(self, flat):
        return [flat[o:o + n].copy() for o, n in zip(self.row_off, self.rows)]

    def solve_device(self, bd: DeviceArray, xd: DeviceArray, tol=1e-6, maxit=10000, want_stats=True):
        it = np.zeros(len(self.rows), np.int32)
        rel = np.zeros(len(self.rows))
        check(lib().ispm_batch_solve(self.b, bd.vp, xd.vp, tol, maxit,
                                     it.ctypes.data if want_stats else None,
                                     rel.ctypes.data if want_stats else None))
        return it, rel

    def __del__(self):
        try:
            lib().ispm_batch_destroy(self.b)
        except Exception:
            pass


def batch_solve(systems, rhs, tol=1e-6, maxit=10000):
    bt = Batch(systems)
    bd = DeviceArray.from_numpy(bt.pack(rhs))
    xd = DeviceArray(bt.rows_padded, np.float32)
    it, rel = bt.solve_device(bd, xd, tol, maxit)
    return bt.unpack(xd.numpy()), it, rel
