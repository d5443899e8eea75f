"""Device-resident FSAI-preconditioned CG (ispm_solver_* in include/ispm.h)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import (BF16, CSR, F32, F64, PLAN_DEFAULT, PLAN_MIN_BYTES, SPMV_PLAN_SPLIT_DIAG, SPMV_PLAN_SYMMETRIC,
               DeviceArray, IspmError, Plan, check, csr_multiply, lib)

PPCG, PCG, PCG1 = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1


class SolverOpts(C.Structure):
    _fields_ = [("method", C.c_int), ("precision", C.c_int), ("tol", C.c_double),
                ("maxit", C.c_int), ("check_every", C.c_int), ("use_graph", C.c_int),
                ("reliable_delta", C.c_double)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("replacements", C.c_int), ("converged", C.c_int),
                ("relres", C.c_double), ("l1_resnorm", C.c_double)]


class NormStats(C.Structure):
    """ispm_norm_stats: the solve with OpenFOAM's normalised residual."""
    _fields_ = [("iterations", C.c_int), ("replacements", C.c_int), ("converged", C.c_int),
                ("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("norm_factor", C.c_double), ("l1_resnorm", C.c_double)]


def fp32_exact(A: CSR) -> bool:
    v = np.asarray(A.vals)
    return v.dtype == np.float32 or bool(np.all(v.astype(np.float32).astype(np.float64) == v))


def _bind():
    L = lib()
    L.ispm_solver_create.restype = C.c_int
    L.ispm_solver_create.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.POINTER(SolverOpts)]
    L.ispm_solver_solve.restype = C.c_int
    L.ispm_solver_solve.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(SolveStats)]
    L.ispm_solver_destroy.restype = None
    L.ispm_solver_destroy.argtypes = [C.c_void_p]
    L.ispm_solver_set_residual_plan.restype = C.c_int
    L.ispm_solver_set_residual_plan.argtypes = [C.c_void_p, C.c_void_p]
    return L


def preconditioner_matrices(M, product=None):
    """The matrices the solver applies for M: (Ainv1, Ainv2) as given, or the one matrix
    Ainv2 * Ainv1 when `product` is True, or (None) when it has fewer than 90% of the two
    factors' nonzeros together (one sparse product streams fewer bytes than two)."""
    if M is None or isinstance(M, CSR):
        return M if M is None else (M,)
    a1, a2 = M
    if product is False:
        return (a1, a2)
    P = csr_multiply(a2, a1, F64)
    if product or P.nnz < 0.9 * (a1.nnz + a2.nnz):
        return (P,)
    return (a1, a2)


def apply_seconds(plans, vt, n, reps=9):
    """Median time of applying the plans in turn (device events), with vectors of the plans'
    own type (vt must be it)."""
    from . import Event, synchronize
    if any(p.vec_type != vt for p in plans):
        raise ValueError("apply_seconds: plans of another vector type")
    dt = np.float64 if vt == F64 else np.float32
    bufs = [DeviceArray.from_numpy(np.ones(n, dt)), DeviceArray(n, dt), DeviceArray(n, dt)]

    def run():
        src = 0
        for k, p in enumerate(plans):
            dst = 1 + (k & 1)
            p.execute(bufs[src], bufs[dst])
            src = dst
    run()
    synchronize()
    e0, e1, ts = Event(), Event(), []
    for _ in range(reps):
        e0.record()
        run()
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_ms(e1))
    return float(np.median(ts))


class Solver:
    """Owns the plans of A and of the preconditioner and a device solver object.

    precision="fp64": A in FP64, preconditioner in FP64, FP64 vectors.  precision="mixed": A
    in FP32, preconditioner in BF16 (factor_type), FP32 vectors, FP64 scalars, reliable FP64
    residual replacement.  M: (Ainv1, Ainv2) from fastainv_sym, applied as Ainv2 (Ainv1 r) or
    as their product (see preconditioner_matrices; fsai_product chooses), or one CSR matrix;
    each matrix as a SlELL plan or a split-diagonal plan (split_diag chooses; None: the
    faster, timed at set-up).  A symmetric matrix on few diagonals may use symmetric diagonal
    storage for A (symmetric chooses; None: the faster, timed at set-up).  A_kind: how A's
    SlELL plan is made (PLAN_DEFAULT, the reference's guess, unless given; PLAN_EXHAUSTIVE
    times the candidate formats, plan.h:97-144)."""

    def __init__(self, A: CSR, M=None, precision="fp64", method=None, tol=1e-10, maxit=100000,
                 check_every=16, use_graph=True, reliable_delta=0.01, factor_type=None,
                 A_plan=None, M_plans=None, fsai_product=None, split_diag=None, symmetric=None,
                 A_kind=None):
        L = _bind()
        mixed = precision == "mixed"
        vt = F32 if mixed else F64
        if method is None:
            method = PCG if mixed else PPCG
        self.A_plan = A_plan
        if A_plan is None:
            # A as a SlELL plan, or (symmetric stencil-like matrices) in symmetric diagonal
            # storage that streams half the matrix: the faster of the two with the solve's
            # vectors (the mixed solve's FP64 residuals use the same plan with FP64 vectors)
            ea = F32 if mixed else F64
            cands = [] if symmetric else [Plan(A, vt, elm_type=ea, kind=PLAN_DEFAULT if A_kind is None else A_kind)]
            if symmetric is not False:
                try:
                    cands.append(Plan(A, vt, elm_type=ea, flags=SPMV_PLAN_SYMMETRIC))
                except IspmError:
                    if symmetric:
                        raise
            self.A_plan = cands[0]
            if len(cands) > 1:
                n = cands[0].n_rows_padded
                t = [apply_seconds((c,), vt, n) for c in cands]
                self.A_plan = cands[int(np.argmin(t))]
        self.M_plans = None
        if M_plans is not None:
            self.M_plans = M_plans
        elif M is not None:
            ft = factor_type if factor_type is not None else (BF16 if mixed else F64)
            # preconditioner rows are irregular: take the byte-minimal format, and also the
            # split-diagonal form (diagonal vector + the rows with off-diagonal entries)
            def plans(mats, split=False):
                if split:
                    return tuple(Plan(m, vt, elm_type=ft, flags=SPMV_PLAN_SPLIT_DIAG) for m in mats)
                return tuple(Plan(m, vt, elm_type=ft, kind=PLAN_MIN_BYTES) for m in mats)
            n = self.A_plan.n_rows_padded
            if fsai_product is None and not isinstance(M, CSR):
                # two factors or their product, each as SlELL or split: whichever applies
                # fastest here (the product has fewer nonzeros but its off-diagonal entries
                # gather x from further away)
                a1, a2 = M
                P = csr_multiply(a2, a1, F64)
                forms = [(a1, a2)] + ([(P,)] if P.nnz < a1.nnz + a2.nnz else [])
            else:
                forms = [preconditioner_matrices(M, fsai_product)]
            best = None
            for mats in forms:
                for split in ((False, True) if split_diag is None else (split_diag,)):
                    cand = plans(mats, split)
                    t = apply_seconds(cand, vt, n) if len(forms) > 1 or split_diag is None else 0.0
                    if best is None or t < best[0]:
                        best = (t, cand)
            self.M_plans = best[1]
        self.n = self.A_plan.n_rows_padded
        o = SolverOpts(method, PREC_MIXED if mixed else PREC_FP64, tol, maxit, check_every,
                       int(use_graph), reliable_delta)
        p = C.c_void_p()
        m1 = self.M_plans[0].p if self.M_plans else None
        m2 = self.M_plans[1].p if self.M_plans and len(self.M_plans) > 1 else None
        check(L.ispm_solver_create(C.byref(p), self.A_plan.p, m1, m2, C.byref(o)))
        self.p = p
        self.res_plan = None
        if mixed and A_plan is None and not fp32_exact(A):
            # FP32 storage changes A: the FP64 true residual needs the FP64 values
            self.res_plan = Plan(A, F64, elm_type=F64)
            check(L.ispm_solver_set_residual_plan(self.p, self.res_plan.p))

    def solve_device(self, b: DeviceArray, x: DeviceArray, allow_fail=False):
        for a in (b, x):
            if a.dtype != np.float64 or a.n < self.n:
                raise ValueError(f"b and x must be FP64 device arrays of at least {self.n} elements")
        st = SolveStats()
        code = lib().ispm_solver_solve(self.p, b.vp, x.vp, C.byref(st))
        if code != 0 and not (allow_fail and code == 6):
            check(code)
        return st

    def solve_normalised(self, b: DeviceArray, x: DeviceArray, tolerance: float, rel_tol: float = 0.0,
                         allow_fail=False) -> NormStats:
        """The solve with OpenFOAM's convergence test (ispm_solver_solve_normalised): residual
        sum|b - A x| / normFactor below `tolerance`, or below `rel_tol` times the initial one."""
        for a in (b, x):
            if a.dtype != np.float64 or a.n < self.n:
                raise ValueError(f"b and x must be FP64 device arrays of at least {self.n} elements")
        st = NormStats()
        code = lib().ispm_solver_solve_normalised(self.p, b.vp, x.vp, float(tolerance), float(rel_tol),
                                                  C.byref(st))
        if code != 0 and not (allow_fail and code == 6):
            check(code)
        return st

    def __del__(self):
        try:
            lib().ispm_solver_destroy(self.p)
        except Exception:
            pass


@dataclass
class Result:
    x: np.ndarray
    iterations: int
    replacements: int
    relres: float
    converged: bool
    l1_resnorm: float


def solve(A: CSR, b, M=None, tol=1e-10, precision="fp64", method=None, x0=None, **kw) -> Result:
    S = Solver(A, M=M, precision=precision, method=method, tol=tol, **kw)
    n = S.n
    bd = DeviceArray.from_numpy(np.concatenate([np.asarray(b, np.float64), np.zeros(n - A.n_rows)]))
    x = np.zeros(n)
    if x0 is not None:
        x[:A.n_rows] = x0
    xd = DeviceArray.from_numpy(x)
    st = S.solve_device(bd, xd, allow_fail=True)
    return Result(xd.numpy()[:A.n_rows], st.iterations, st.replacements, st.relres,
                  bool(st.converged), st.l1_resnorm)
