"""The call pattern of the reference library's caller, the foam-extend ispm linear solver
(README.md:3-9 of the reference; the caller itself is not part of the reference): an OpenFOAM
lduMatrix (a diagonal coefficient per cell, upper and lower coefficients per internal face
with owner < neighbour addressing) converted to CSR once; FSAI-preconditioned CG stopped on
OpenFOAM's normalised residual (tolerance / relTol); the plan of A kept across time steps,
with each step's coefficients scattered into it on the device; the drop tolerance of the
approximate inverse searched across time steps by DroptolTuner
(fastainv/droptol-tuning.h:4-103), which is fed each step's set-up + solve time.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import (CSR, F32, F64, SPMV_PLAN_SPLIT_DIAG, SPMV_PLAN_SYMMETRIC, CsrOut, DeviceArray, IspmError,
               Plan, _csr_from_out, check, fastainv_sym, lib, synchronize)
from .solver import PCG, PPCG, Solver


@dataclass
class LduMatrix:
    """OpenFOAM lduMatrix coefficients and addressing: face f joins cells lower_addr[f] <
    upper_addr[f]; upper[f] = A[lower][upper], lower[f] = A[upper][lower] (None: symmetric)."""
    n: int
    lower_addr: np.ndarray
    upper_addr: np.ndarray
    diag: np.ndarray
    upper: np.ndarray
    lower: Optional[np.ndarray] = None

    @property
    def n_faces(self):
        return int(len(self.lower_addr))

    def arrays(self):
        la = np.ascontiguousarray(self.lower_addr, dtype=np.int32)
        ua = np.ascontiguousarray(self.upper_addr, dtype=np.int32)
        d = np.ascontiguousarray(self.diag, dtype=np.float64)
        u = np.ascontiguousarray(self.upper, dtype=np.float64)
        lo = None if self.lower is None else np.ascontiguousarray(self.lower, dtype=np.float64)
        if len(d) != self.n or len(ua) != len(la) or len(u) != len(la) or (lo is not None and len(lo) != len(la)):
            raise ValueError("lduMatrix arrays of inconsistent lengths")
        return la, ua, d, u, lo


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def ldu_to_csr(ldu: LduMatrix):
    """(CSR, pos): the CSR of the lduMatrix (ispm_ldu_to_csr) and the CSR index of every
    coefficient (diag, then upper, then lower)."""
    la, ua, d, u, lo = ldu.arrays()
    pos = np.empty(ldu.n + 2 * ldu.n_faces, np.int32)
    o = CsrOut()
    check(lib().ispm_ldu_to_csr(ldu.n, ldu.n_faces, _p(la), _p(ua), _p(d), _p(u), _p(lo), C.byref(o), _p(pos)))
    try:
        return _csr_from_out(o), pos
    finally:
        lib().ispm_csr_out_free(C.byref(o))


def ldu_values(ldu: LduMatrix, pos: np.ndarray) -> np.ndarray:
    """The coefficients in CSR value order (ispm_ldu_values)."""
    la, ua, d, u, lo = ldu.arrays()
    vals = np.empty(ldu.n + 2 * ldu.n_faces, np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    check(lib().ispm_ldu_values(ldu.n, ldu.n_faces, _p(pos), _p(d), _p(u), _p(lo), _p(vals)))
    return vals


class DroptolTuner:
    """fastainv/droptol-tuning.h's DroptolTuner through the C-ABI: probe droptol(), report
    note_time(t), read best()."""

    def __init__(self, log2_lo=-10.0, log2_hi=-3.0):
        p = C.c_void_p()
        check(lib().ispm_droptol_tuner_create(float(log2_lo), float(log2_hi), C.byref(p)))
        self.p = p

    def droptol(self) -> float:
        return lib().ispm_droptol_tuner_droptol(self.p)

    def best(self) -> float:
        return lib().ispm_droptol_tuner_best(self.p)

    def note_time(self, seconds: float):
        lib().ispm_droptol_tuner_note_time(self.p, float(seconds))

    def restart(self):
        lib().ispm_droptol_tuner_restart(self.p)

    def iterations(self) -> int:
        return lib().ispm_droptol_tuner_iterations(self.p)

    def __del__(self):
        try:
            lib().ispm_droptol_tuner_destroy(self.p)
        except Exception:
            pass


@dataclass
class SolverPerformance:
    """What an OpenFOAM solver reports for one solve, plus the set-up figures (setup_seconds
    includes fsai_seconds, the host computation of the approximate inverse)."""
    initial_residual: float
    final_residual: float
    n_iterations: int
    converged: bool
    droptol: float
    setup_seconds: float
    solve_seconds: float
    fsai_seconds: float = 0.0


class FoamPCG:
    """FSAI-PCG for a sequence of lduMatrix systems with one addressing (time steps).

    The A plan is built once and its values are replaced by the device refresh: a symmetric
    diagonal plan when the lduMatrix is symmetric and its faces lie on at most 16 diagonals
    (structured meshes: half the bytes of the general layout), else a SlELL plan;
    `update(ldu)` scatters a new step's coefficients into it on the device (and into the host
    CSR the approximate inverse is computed from).  `solve(source, psi, tolerance, rel_tol)`
    computes the approximate inverse at the tuner's drop tolerance, solves on the device with
    OpenFOAM's normalised residual and reports the set-up + solve time to the tuner.
    precision "fp64" or "mixed" (FP32/BF16 iterations, FP64 residuals).

    The preconditioner's storage (the two factors or their product, each as SlELL or
    split-diagonal plans) is chosen by timing the candidates at the first solve and kept for
    the following ones (the addressing is fixed); `redecide_every` > 0 times them again every
    that many solves."""

    def __init__(self, ldu: LduMatrix, precision="fp64", method=None, droptol=None,
                 log2_range=(-10.0, -3.0), max_elts_per_row=64, maxit=1000, symmetric=None,
                 redecide_every=0):
        self.precision = precision
        self.mixed = precision == "mixed"
        self.method = method if method is not None else (PCG if self.mixed else PPCG)
        self.maxit = maxit
        self.max_elts_per_row = max_elts_per_row
        self.csr, self.pos = ldu_to_csr(ldu)
        self.n, self.n_faces = ldu.n, ldu.n_faces
        vt = F32 if self.mixed else F64
        et = F32 if self.mixed else F64
        # symmetric=None: symmetric storage where the matrix allows it
        if symmetric is None:
            symmetric = ldu.lower is None
        elif symmetric and ldu.lower is not None:
            raise ValueError("symmetric storage needs a symmetric lduMatrix (lower None)")
        self.A_plan = None
        if symmetric:
            try:
                self.A_plan = Plan(self.csr, vt, elm_type=et, flags=SPMV_PLAN_SYMMETRIC)
            except IspmError:
                pass   # more than 16 diagonals: the general layout
        if self.A_plan is None:
            self.A_plan = Plan(self.csr, vt, elm_type=et)
        flags = self.A_plan.flags & SPMV_PLAN_SYMMETRIC
        # the FP64 true residual reads A in FP64 (FP32 storage rounds the coefficients)
        self.res_plan = Plan(self.csr, F64, elm_type=F64, flags=flags) if self.mixed else None
        self.n_pad = self.A_plan.n_rows_padded
        self.d_pos = DeviceArray.from_numpy(self.pos)
        self.d_vals = DeviceArray(self.csr.nnz, np.float64)
        self.fixed_droptol = droptol
        self.tuner = None if droptol is not None else DroptolTuner(*log2_range)
        self.steps = 0
        self.redecide_every = int(redecide_every)
        self.storage = None      # (fsai_product, split_diag) chosen at the last decision

    def update(self, ldu: LduMatrix):
        """New coefficients, same addressing: refresh the plan(s) of A on the device."""
        if ldu.n != self.n or ldu.n_faces != self.n_faces:
            raise ValueError("update: the addressing changed (build a new FoamPCG)")
        la, ua, d, u, lo = ldu.arrays()
        dd, du = DeviceArray.from_numpy(d), DeviceArray.from_numpy(u)
        dl = DeviceArray.from_numpy(lo) if lo is not None else None
        check(lib().ispm_ldu_values_device(self.n, self.n_faces, self.d_pos.vp, dd.vp, du.vp,
                                           dl.vp if dl is not None else None, self.d_vals.vp))
        self.A_plan.update_coefs_device(self.d_vals)
        if self.res_plan is not None:
            self.res_plan.update_coefs_device(self.d_vals)
        self.csr = CSR(self.csr.n_rows, self.csr.n_cols, self.csr.row_ptr, self.csr.cols,
                       ldu_values(ldu, self.pos))
        synchronize()

    def droptol(self) -> float:
        return self.fixed_droptol if self.tuner is None else self.tuner.droptol()

    def solve(self, source, psi, tolerance=1e-6, rel_tol=0.0):
        """(psi, SolverPerformance): one solve of A psi = source from the given psi."""
        source = np.asarray(source, np.float64)
        psi = np.asarray(psi, np.float64)
        if source.shape != (self.n,) or psi.shape != (self.n,):
            raise ValueError(f"source and psi must have {self.n} entries")
        dt = self.droptol()
        synchronize()
        t0 = time.perf_counter()
        ft = F32 if self.mixed else F64
        factors = fastainv_sym(self.csr, ft, max_elts=self.max_elts_per_row * self.n, droptol=dt)
        tf = time.perf_counter()
        decide = self.storage is None or (self.redecide_every > 0 and self.steps % self.redecide_every == 0)
        product, split = (None, None) if decide else self.storage
        solver = Solver(self.csr, M=factors, precision=self.precision, method=self.method,
                        A_plan=self.A_plan, maxit=self.maxit, fsai_product=product, split_diag=split)
        if decide and solver.M_plans:
            self.storage = (len(solver.M_plans) == 1,
                            bool(solver.M_plans[0].flags & SPMV_PLAN_SPLIT_DIAG))
        if self.res_plan is not None:
            check(lib().ispm_solver_set_residual_plan(solver.p, self.res_plan.p))
        pad = np.zeros(self.n_pad - self.n)
        b = DeviceArray.from_numpy(np.concatenate([source, pad]))
        x = DeviceArray.from_numpy(np.concatenate([psi, pad]))
        synchronize()
        t1 = time.perf_counter()
        st = solver.solve_normalised(b, x, tolerance, rel_tol, allow_fail=True)
        synchronize()
        t2 = time.perf_counter()
        if self.tuner is not None:
            self.tuner.note_time(t2 - t0)
        self.steps += 1
        perf = SolverPerformance(st.initial_residual, st.final_residual, st.iterations, bool(st.converged),
                                 dt, t1 - t0, t2 - t1, tf - t0)
        return x.numpy(self.n), perf
