"""Benchmark: FSAI-PCG time to a 1e-10 FP64 relative residual (mixed precision vs FP64) and
SlELL SpMV HBM bandwidth on one MI355X.

Workload: 27-point stencil on a 192^3 grid (7,077,888 rows, 189,119,224 nonzeros), coefficient
contrast 1e4 in random 8^3 blocks, synthetic and seeded.  A "step" is one complete solve of
A x = b (b = A x_true, x_true ~ U[-1,1]) from x = 0 to ||b - A x||_2 / ||b||_2 <= 1e-10, with
the matrix, the preconditioner factors and b already resident in HBM.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ispm-sparse-lib_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class CacheFlush:
    """Evicts the 256 MB Infinity Cache between timed launches: a device-to-device copy of
    512 MiB (1 GiB of traffic) on the library stream, outside the timed events."""

    def __init__(self, I, nbytes=512 << 20):
        self.I, self.n = I, nbytes
        self.src, self.dst = I.DeviceArray(nbytes // 8), I.DeviceArray(nbytes // 8)

    def __call__(self):
        self.I.check(self.I.lib().ispm_memcpy_async(self.dst.vp, self.src.vp, self.n, 2))   # ISPM_D2D


def time_spmv(I, run, steps, warmup, flush=None):
    """(average, median) launch duration of one SpMV (`run()` launches it), timed with HIP
    events recorded on the library stream (the stream the kernel runs on).  Without `flush`,
    five batches of `steps` back-to-back launches, each batch between two events (the
    launches' own time, not the event overhead of single launches); with `flush`, the cache is
    flushed before every timed launch (for matrices that fit in the Infinity Cache), each
    launch between its own two events."""
    for _ in range(warmup):
        run()
    I.synchronize()
    e0, e1 = I.Event(), I.Event()
    ts = []
    if flush:
        for _ in range(steps):
            flush()
            e0.record()
            run()
            e1.record()
            e1.synchronize()
            ts.append(e0.elapsed_ms(e1) * 1e-3)
    else:
        for _ in range(5):
            e0.record()
            for _ in range(steps):
                run()
            e1.record()
            e1.synchronize()
            ts.append(e0.elapsed_ms(e1) * 1e-3 / steps)
    return float(np.mean(ts)), float(np.median(ts))


def spmv_leg(I, G, A, vt, et, steps, warmup, flags=0, flush=None, guarded=False, kind=None):
    """One plan's SpMV timed; `guarded`: x is held with the plan's zero guard zones and the
    product runs through ispm_plan_execute_guarded, as the solver's own vectors are; `kind`:
    the planning mode (default: the reference's default plan)."""
    plan = I.Plan(A, vt, elm_type=et, flags=flags, kind=I.PLAN_DEFAULT if kind is None else kind)
    inf = plan.info()
    n = plan.n_rows_padded
    dt = np.float64 if vt == I.F64 else np.float32
    g = plan.x_guard() if guarded else 0
    xh = np.zeros(n + 2 * g, dt)
    xh[g:g + A.n_rows] = G.uniform(0, A.n_rows)
    x = I.DeviceArray.from_numpy(xh)
    y = I.DeviceArray(n, dt)
    run = (lambda: plan.execute_guarded(x, g, y)) if guarded else (lambda: plan.execute(x, y))
    mean, med = time_spmv(I, run, steps, warmup, flush)
    stored = inf.stored_bytes
    return {
        "format": dict(S=inf.S, hblock=inf.hblock, V=inf.V, D=inf.D,
                       N=inf.launch.N, H=inf.launch.H,
                       symmetric=bool(flags & I.SPMV_PLAN_SYMMETRIC)),
        "avg_s": mean, "median_s": med,
        "stored_bytes": stored, "csr_bytes": inf.csr_bytes,
        "GBps": stored / mean * 1e-9, "csr_equiv_GBps": inf.csr_bytes / mean * 1e-9,
        "GFLOPs": 2.0 * A.nnz / mean * 1e-9,
        "frac_hbm": stored / mean * 1e-9 / HBM_PEAK_GBS,
        "cache_flushed": flush is not None,
        "x_guard": g,
    }


def roofline(leg, key, kernel):
    """HBM roofline entry of one SpMV leg: achieved = stored (algorithmic) bytes per launch over
    the live average launch duration; traffic = the HBM bytes per launch that rocprofv3
    FETCH_SIZE / WRITE_SIZE measured for the same plan (profiles/spmv_27pt_192_traffic.json,
    counters scaled by the factors a 1 GiB copy gave), null when no measurement of this
    format is on record."""
    traffic = None
    path = os.path.join(ROOT, "profiles", "spmv_27pt_192_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            rec = json.load(f).get(key)
        same = rec and rec.get("stored_bytes") == leg["stored_bytes"] and all(
            rec["format"].get(k) == leg["format"][k] for k in ("S", "hblock", "V", "D")) and \
            bool(rec["format"].get("symmetric", False)) == leg["format"]["symmetric"] and \
            bool(rec.get("x_guarded", False)) == (leg["x_guard"] > 0)
        if same:
            traffic = rec["traffic_bytes"]
    return {"bound": "hbm", "kernel": kernel, "achieved": leg["GBps"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": leg["frac_hbm"], "traffic": traffic,
            "algorithmic_bytes": leg["stored_bytes"]}


def host_spmv(A, x):
    """A x in FP64 on the host (scipy), to make right-hand sides of known solutions."""
    import scipy.sparse as sp
    M = sp.csr_matrix((np.asarray(A.vals, np.float64), A.cols, A.row_ptr), shape=(A.n_rows, A.n_cols))
    return M @ np.asarray(x, np.float64)


def cpu_spmv_baseline(O, A, seconds=10.0):
    threads = O.set_threads(0)
    x = np.random.default_rng(0).uniform(-1, 1, A.n_rows)
    O.csr_spmv(A, x)
    reps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds and reps < 50:
        O.csr_spmv(A, x)
        reps += 1
    dt = (time.perf_counter() - t0) / reps
    csr = 4 * (A.n_rows + 1) + 12 * A.nnz + 16 * A.n_rows
    return {"value": csr / dt * 1e-9, "unit": "GB/s (CSR-equivalent, FP64 SpMV)", "cores": threads,
            "kind": "port", "sample": f"{reps} FP64 CSR SpMVs of the full 192^3 27-point matrix",
            "seconds_per_spmv": dt}


def solve_leg(I, S, O, A, b, factors, precision, method, steps, warmup, tol, **solver_kw):
    """Time `steps` complete solves (after `warmup` untimed ones) from x = 0; every solve is
    checked with the CPU oracle's FP64 CSR residual."""
    solver = S.Solver(A, M=factors, precision=precision, method=method, tol=tol, **solver_kw)
    n = solver.n
    bd = I.DeviceArray.from_numpy(np.concatenate([b, np.zeros(n - A.n_rows)]))
    xd = I.DeviceArray(n)
    times, stats = [], None
    for it in range(warmup + steps):
        xd.zero()
        I.synchronize()
        t0 = time.perf_counter()
        stats = solver.solve_device(bd, xd, allow_fail=True)
        I.synchronize()
        dt = time.perf_counter() - t0
        if it >= warmup:
            times.append(dt)
    x = xd.numpy()[:A.n_rows]
    r = b - O.csr_spmv(A, x)
    rel = float(np.linalg.norm(r) / np.linalg.norm(b))
    pre = solver.M_plans or ()
    return {"precision": precision,
            "method": {0: "pipelined-pcg", 1: "pcg", 2: "single-reduction-pcg"}[method],
            "A_storage": ("symmetric diagonal planes" if solver.A_plan.flags & I.SPMV_PLAN_SYMMETRIC
                          else "SlELL"),
            "preconditioner": ("product Ainv2*Ainv1" if len(pre) == 1 else "factors Ainv1, Ainv2")
                              + (", split-diagonal plans" if pre and pre[0].flags & I.SPMV_PLAN_SPLIT_DIAG
                                 else ", SlELL plans"),
            "preconditioner_nnz_per_row": sum(p.info().device.n_nz for p in pre) / A.n_rows,
            "seconds": float(np.mean(times)), "seconds_min": float(np.min(times)),
            "iterations": stats.iterations, "replacements": stats.replacements,
            "converged": bool(stats.converged), "gpu_relres": stats.relres,
            "oracle_relres": rel}, x


def cpu_solve_baseline(O, A, b, tol, max_reps=5, budget_s=45.0, droptol=0.125):
    """The CPU path on the host cores: the oracle's FP64 PCG (Hestenes-Stiefel, host CSR
    products over all cores, the method of the fastest GPU FP64 solve) from x = 0 to the same
    FP64 relative residual with the same preconditioner, timed to convergence, median of up
    to `max_reps` solves within `budget_s`; and the restated fastainv_sym (one core: its
    column loop is sequential), timed once, whose factors the solves use."""
    import platform
    t = time.perf_counter()
    f = O.fastainv_sym(A, out_double=True, max_elts=64 * A.n_rows, droptol=droptol)
    fsai_s = time.perf_counter() - t
    M = [type(A)(A.n_rows, A.n_cols, r, c, v) for r, c, v in f]
    threads = O.set_threads(0)
    ts, it, rel = [], 0, None
    t0 = time.perf_counter()
    while len(ts) < max_reps and (not ts or time.perf_counter() - t0 + ts[-1] <= budget_s):
        t = time.perf_counter()
        _, it, rel = O.pcg(A, b, M1=M[0], M2=M[1], tol=tol, maxit=100000)
        ts.append(time.perf_counter() - t)
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            cpu = next(l.split(":", 1)[1].strip() for l in fh if l.startswith("model name"))
    except (OSError, StopIteration):
        cpu = platform.processor() or cpu
    return {"value": float(np.median(ts)), "unit": "s (FP64 FSAI-PCG solve to 1e-10 on the host)",
            "cores": threads, "kind": "port",
            "sample": f"{len(ts)} complete solves of the full matrix from x = 0 (the oracle's FP64 "
                      f"Hestenes-Stiefel PCG, {it} iterations, true relres {rel:.2e}), median; "
                      f"the oracle's fastainv_sym on 1 core: {fsai_s:.2f} s",
            "cpu_model": cpu, "solve_seconds": ts, "iterations": it, "relres": rel,
            "fastainv_seconds_1core": fsai_s}


def foam_legs(I, G, O, m=128, steps=4, tol=1e-8):
    """The foam-extend call pattern on an m^3 diffusion mesh given as an lduMatrix: the CSR
    built once, then `steps` time steps whose new coefficients (lognormal conductances of a
    new seed) are scattered into the same A plan on the device; each step computes the
    approximate inverse at the DroptolTuner's probe and solves (FP64 pipelined PCG) to
    OpenFOAM's normalised residual `tol` from the previous step's field, towards a new field
    every step.  The approximate inverse is computed on the host, as the reference does."""
    from ispm_amd import foam as F
    t = time.perf_counter()
    L = F.LduMatrix(*G.ldu_mesh(m, m, m, sigma=1.0, seed=5001))
    drv = F.FoamPCG(L, precision="fp64")
    log(f"foam {m}^3: {L.n} cells, {L.n_faces} faces, CSR and plan in {time.perf_counter() - t:.1f}s")
    psi = np.zeros(L.n)
    legs = []
    for k in range(steps):
        Lk = F.LduMatrix(*G.ldu_mesh(m, m, m, sigma=1.0, seed=5001 + k))
        t = time.perf_counter()
        if k:
            drv.update(Lk)
        upd = time.perf_counter() - t
        A, _ = F.ldu_to_csr(Lk)
        b = host_spmv(A, G.uniform(5100 + k, L.n))   # a new field to reach every step
        psi0 = psi.copy()
        psi, perf = drv.solve(b, psi0, tolerance=tol)
        nf, r = O.foam_residual(A, b, psi0, psi)
        legs.append({"step": k, "droptol": perf.droptol, "refresh_s": upd, "setup_s": perf.setup_seconds,
                     "fsai_s": perf.fsai_seconds, "solve_s": perf.solve_seconds, "iterations": perf.n_iterations,
                     "initial_residual": perf.initial_residual, "final_residual": perf.final_residual,
                     "oracle_final_residual": r, "converged": perf.converged})
        log(f"foam step {k}: droptol {perf.droptol:.4g}, refresh {upd * 1e3:.1f} ms, set-up "
            f"{perf.setup_seconds:.2f} s ({perf.fsai_seconds:.2f} s approximate inverse), solve {perf.solve_seconds * 1e3:.1f} ms, {perf.n_iterations} it, "
            f"residual {perf.initial_residual:.2e} -> {perf.final_residual:.2e} (oracle {r:.2e})")
    return {"mesh": f"{m}^3", "cells": L.n, "faces": L.n_faces, "tolerance": tol, "steps": legs}


def hblock_legs(I, G, O, steps, warmup):
    """SpMV of a banded two-unknowns-per-node matrix (1,048,576 nodes, 2,097,152 rows, every
    row aligned column pairs within 40 nodes: the format guess takes hblock pairs) with the
    default plan, whose slices read x from LDS windows, against the same device layout through
    the generic slice kernel that gathers x from global memory (the reference's mapping)."""
    import ctypes as C
    A = G.banded_pairs(1 << 20, seed=5)
    x = G.uniform(6, A.n_rows)
    out = {"n_rows": A.n_rows, "nnz": A.nnz}
    for name, vt, et in (("fp64", I.F64, I.F64), ("fp32", I.F32, I.F32)):
        plan = I.Plan(A, vt, elm_type=et)
        inf = plan.info()
        n = plan.n_rows_padded
        dt = np.float64 if vt == I.F64 else np.float32
        xd = I.DeviceArray.from_numpy(np.concatenate([x, np.zeros(n - A.n_rows)]).astype(dt))
        y1, y2 = I.DeviceArray(n, dt), I.DeviceArray(n, dt)
        lp = inf.launch
        t_win, _ = time_spmv(I, lambda: plan.execute(xd, y1), steps, warmup)
        t_glob, _ = time_spmv(I, lambda: I.check(I.lib().ispm_spmv(C.byref(lp), C.byref(inf.device), vt,
                                                                   xd.vp, y2.vp)), steps, warmup)
        I.synchronize()
        same = bool(np.array_equal(y1.numpy(), y2.numpy()))
        ref = O.csr_spmv(A, x.astype(dt).astype(np.float64))
        err = float(np.max(np.abs(y1.numpy()[:A.n_rows] - ref) / np.maximum(O.csr_spmv_abs(A, x), 1e-300)))
        stored = inf.stored_bytes
        out[name] = {"format": dict(S=inf.S, hblock=inf.hblock, V=inf.V, D=inf.D, N=lp.N, H=lp.H),
                     "window_slices": inf.window_slices, "slices": inf.device.n_slices,
                     "stored_bytes": stored, "windows_s": t_win, "global_gathers_s": t_glob,
                     "windows_GBps": stored / t_win * 1e-9, "global_GBps": stored / t_glob * 1e-9,
                     "frac_hbm": stored / t_win * 1e-9 / HBM_PEAK_GBS,
                     "bit_identical_to_global": same, "max_rel_err_vs_oracle": err}
        log(f"hblock {name}: S={inf.S} windows {inf.window_slices}/{inf.device.n_slices} slices, "
            f"{t_win * 1e6:.1f} us ({stored / t_win * 1e-9:.0f} GB/s) vs global gathers "
            f"{t_glob * 1e6:.1f} us ({stored / t_glob * 1e-9:.0f} GB/s); same {same}, rel err {err:.1e}")
        del plan, xd, y1, y2
    return out


def spmm_legs(I, G, O, steps, warmup):
    """Multi-vector products (k = 3 and 8 columns, column-major with leading dimension n) of a
    7-point and a 27-point 128^3 stencil, FP64 and FP32, with the default (diagonal-pattern)
    plan and with symmetric diagonal storage: times against the single-vector product, the
    bytes of the matrix plus k columns read and written, and column k-1 compared bit for bit
    with the plan's single-vector product and against the oracle."""
    out = {}
    for mname, A in (("7pt128", G.poisson7(128)), ("27pt128", G.stencil27_contrast(128))):
        x = G.uniform(7, 8 * A.n_rows)
        for tname, vt in (("fp64", I.F64), ("fp32", I.F32)):
            dt = np.float64 if vt == I.F64 else np.float32
            for sname, fl in (("pattern", 0), ("symmetric", I.SPMV_PLAN_SYMMETRIC)):
                plan = I.Plan(A, vt, elm_type=vt, flags=fl)
                inf = plan.info()
                n = plan.n_rows_padded
                X = np.zeros((8, n), dt)
                X[:, :A.n_rows] = x.reshape(8, A.n_rows)
                Xd = I.DeviceArray.from_numpy(X.ravel())
                Yd = I.DeviceArray(8 * n, dt)
                y1 = I.DeviceArray(n, dt)
                leg = {"format": dict(S=inf.S, D=inf.D, symmetric=bool(fl)), "stored_bytes": inf.stored_bytes,
                       "pattern_planes": inf.pattern_planes}
                t1, _ = time_spmv(I, lambda: plan.execute(Xd, y1), steps, warmup)
                leg["k1_s"] = t1
                vec = (n + A.n_rows) * np.dtype(dt).itemsize
                for k in (3, 8):
                    tk, _ = time_spmv(I, lambda: plan.execute_multi(k, Xd, n, Yd, n), steps, warmup)
                    xk = I.DeviceArray.from_numpy(np.ascontiguousarray(X[k - 1]))
                    plan.execute(xk, y1)
                    I.synchronize()
                    yk = Yd.numpy()[(k - 1) * n:k * n]
                    same = bool(np.array_equal(yk, y1.numpy()))
                    xr = X[k - 1, :A.n_rows].astype(np.float64)
                    err = float(np.max(np.abs(yk[:A.n_rows] - O.csr_spmv(A, xr)) / np.maximum(O.csr_spmv_abs(A, xr), 1e-300)))
                    byts = inf.stored_bytes - vec + k * vec
                    leg[f"k{k}"] = {"seconds": tk, "GBps": byts / tk * 1e-9, "per_column_vs_k1": tk / k / t1,
                                    "column_bit_identical": same, "max_rel_err_vs_oracle": err}
                    log(f"spmm {mname} {tname} {sname} k={k}: {tk * 1e6:.1f} us ({byts / tk * 1e-9:.0f} GB/s, "
                        f"{tk / k / t1:.2f} of a single product per column), same {same}, rel err {err:.1e}")
                out[f"{mname}_{tname}_{sname}"] = leg
                del plan, Xd, Yd, y1
    return out


def batching_legs(I, G, O, S):
    """The other workload configurations, timed on the device with inputs resident: the 128^3
    finite-volume operator (FP64 SpMV and FSAI-PCG to 1e-8), its 3 and 8 right-hand sides as
    one block PCG against as many single-RHS solves with the same preconditioner, and 512
    uneven random-geometric-graph systems batched against the same systems as one
    block-diagonal matrix through the single-system solver."""
    import ctypes as C
    from ispm_amd import batch as B
    out = {}
    t = time.perf_counter()
    A = G.fv7_lognormal(128)
    log(f"generated 128^3 finite-volume: n={A.n_rows} nnz={A.nnz} in {time.perf_counter() - t:.1f}s")
    flush = CacheFlush(I)
    out["cfg2_spmv_fp64"] = spmv_leg(I, G, A, I.F64, I.F64, 20, 3, flush=flush)
    # the same product with the plan the exhaustive autotune picks (plan.h:97-144)
    out["cfg2_spmv_fp64_tuned"] = spmv_leg(I, G, A, I.F64, I.F64, 20, 3, flush=flush, kind=I.PLAN_EXHAUSTIVE)
    log(f"cfg2 spmv: default plan {out['cfg2_spmv_fp64']['avg_s'] * 1e6:.1f} us, tuned "
        f"{out['cfg2_spmv_fp64_tuned']['avg_s'] * 1e6:.1f} us (cache flushed)")
    M = I.fastainv_sym(A, I.F64, max_elts=64 * A.n_rows)
    b = host_spmv(A, G.uniform(1, A.n_rows))
    sol, _ = solve_leg(I, S, O, A, b, M, "fp64", S.PCG, 3, 1, 1e-8)
    out["cfg2_pcg_fp64"] = sol
    sol, _ = solve_leg(I, S, O, A, b, M, "fp64", S.PCG, 3, 1, 1e-8, A_kind=I.PLAN_EXHAUSTIVE)
    out["cfg2_pcg_fp64_tuned"] = sol
    log(f"cfg2 pcg: {out['cfg2_pcg_fp64']['seconds'] * 1e3:.1f} ms, {sol['iterations']} it; with the "
        f"autotuned A plan {sol['seconds'] * 1e3:.1f} ms")
    # cfg4a: block PCG (SpMM over k columns, the two factors as SlELL plans) vs k single solves,
    # both with A's plan from the exhaustive autotune
    Ap = I.Plan(A, I.F64, elm_type=I.F64, kind=I.PLAN_EXHAUSTIVE)
    Mp = (I.Plan(M[0], I.F64, elm_type=I.F64, kind=I.PLAN_MIN_BYTES),
          I.Plan(M[1], I.F64, elm_type=I.F64, kind=I.PLAN_MIN_BYTES))
    n = Ap.n_rows_padded
    single = S.Solver(A, M_plans=Mp, A_plan=Ap, precision="fp64", method=S.PCG, tol=1e-8)
    for k in (3, 8):
        Bm = np.zeros((k, n))
        for j in range(k):
            Bm[j, :A.n_rows] = host_spmv(A, G.uniform(300 + j, A.n_rows))
        opts = S.SolverOpts(S.PCG, S.PREC_FP64, 1e-8, 100000, 16, 1, 0.1)   # graph-replayed batches
        bs = C.c_void_p()
        I.check(I.lib().ispm_block_solver_create(C.byref(bs), Ap.p, Mp[0].p, Mp[1].p, k, C.byref(opts)))
        bd = I.DeviceArray.from_numpy(Bm.ravel())
        xd = I.DeviceArray(k * n)
        stats = (S.SolveStats * k)()
        ts = []
        for rep in range(3):
            xd.zero()
            I.synchronize()
            t0 = time.perf_counter()
            I.check(I.lib().ispm_block_solver_solve(bs, bd.vp, n, xd.vp, n, stats))
            I.synchronize()
            ts.append(time.perf_counter() - t0)
        X = xd.numpy().reshape(k, n)[:, :A.n_rows]
        worst = max(float(np.linalg.norm(Bm[j, :A.n_rows] - O.csr_spmv(A, X[j])) / np.linalg.norm(Bm[j, :A.n_rows]))
                    for j in range(k))
        I.lib().ispm_block_solver_destroy(bs)
        t1 = []
        for j in range(k):
            bj = I.DeviceArray.from_numpy(Bm[j].copy())
            xj = I.DeviceArray(n)
            single.solve_device(bj, xj)   # warm-up (graph capture)
            xj.zero()
            I.synchronize()
            t0 = time.perf_counter()
            single.solve_device(bj, xj)
            I.synchronize()
            t1.append(time.perf_counter() - t0)
        out[f"cfg4a_block_k{k}"] = {"seconds": float(np.min(ts)), "iterations": [st.iterations for st in stats],
                                    "oracle_relres_max": worst, "single_rhs_seconds_sum": float(np.sum(t1)),
                                    "speedup_vs_single": float(np.sum(t1) / np.min(ts))}
        log(f"cfg4a k={k}: block {np.min(ts) * 1e3:.1f} ms vs {np.sum(t1) * 1e3:.1f} ms single solves")
    del single, Ap, Mp, A, M
    # cfg4b: 512 uneven systems in one persistent launch vs one launch per system
    count = 512
    sizes = np.exp(np.log(1000) + (np.log(64000) - np.log(1000)) * G.uniform01(4001, count)).astype(int)
    ks = 10 + np.minimum(16, (G.uniform01(4002, count) * 17).astype(int))
    t = time.perf_counter()
    systems = [G.rgg_laplacian(int(m), int(kk), seed=5000 + i, dtype=np.float32)
               for i, (m, kk) in enumerate(zip(sizes, ks))]
    rhs = [host_spmv(Ai, G.uniform(6000 + i, Ai.n_rows)).astype(np.float32) for i, Ai in enumerate(systems)]
    log(f"generated 512 systems ({sum(a.n_rows for a in systems)} rows) in {time.perf_counter() - t:.1f}s")
    bt = B.Batch(systems)
    bd = I.DeviceArray.from_numpy(bt.pack(rhs))
    xd = I.DeviceArray(bt.rows_padded, np.float32)
    ts = []
    for rep in range(3):
        xd.zero()
        I.synchronize()
        t0 = time.perf_counter()
        it, rel = bt.solve_device(bd, xd, 1e-6, 20000, want_stats=False)
        I.synchronize()
        ts.append(time.perf_counter() - t0)
    it, rel = bt.solve_device(bd, xd, 1e-6, 20000)
    xs = bt.unpack(xd.numpy())
    worst = 0.0
    for Ai, bi, xi in zip(systems, rhs, xs):
        A64 = I.CSR(Ai.n_rows, Ai.n_cols, Ai.row_ptr, Ai.cols, Ai.vals.astype(np.float64))
        r = bi.astype(np.float64) - O.csr_spmv(A64, xi.astype(np.float64))
        worst = max(worst, float(np.linalg.norm(r) / np.linalg.norm(bi)))
    # baseline: the same systems as one block-diagonal matrix through the single-system
    # solver (Jacobi preconditioner as the diagonal factor pair, mixed precision, PCG), to the
    # same relative tolerance on the whole right-hand side
    offs = np.cumsum([0] + [a.n_rows for a in systems])
    N = int(offs[-1])
    rp = np.zeros(N + 1, np.int64)
    nz = np.cumsum([0] + [a.nnz for a in systems])
    for i, a in enumerate(systems):
        rp[offs[i] + 1:offs[i + 1] + 1] = a.row_ptr[1:].astype(np.int64) + nz[i]
    cols = np.concatenate([a.cols.astype(np.int64) + offs[i] for i, a in enumerate(systems)]).astype(np.int32)
    vals = np.concatenate([a.vals.astype(np.float64) for a in systems])
    Abd = I.CSR(N, N, rp.astype(np.int32), cols, vals)
    rows_bd = np.repeat(np.arange(N), np.diff(rp))
    dg = np.zeros(N)
    np.add.at(dg, rows_bd[cols == rows_bd], vals[cols == rows_bd])
    Dm = I.CSR(N, N, np.arange(N + 1, dtype=np.int32), np.arange(N, dtype=np.int32), 1.0 / np.sqrt(dg))
    bb = np.concatenate([r.astype(np.float64) for r in rhs])
    t = time.perf_counter()
    sbd = S.Solver(Abd, M=(Dm, Dm), precision="mixed", method=S.PCG, tol=1e-6, factor_type=I.F32)
    log(f"block-diagonal baseline: n={N} nnz={Abd.nnz}, solver set up in {time.perf_counter() - t:.1f}s")
    nb = sbd.n
    bdd = I.DeviceArray.from_numpy(np.concatenate([bb, np.zeros(nb - N)]))
    xbd = I.DeviceArray(nb)
    tb = []
    for rep in range(3):
        xbd.zero()
        I.synchronize()
        t0 = time.perf_counter()
        stb = sbd.solve_device(bdd, xbd, allow_fail=True)
        I.synchronize()
        tb.append(time.perf_counter() - t0)
    xb = xbd.numpy()[:N]
    worst_bd = 0.0
    for i, (Ai, bi) in enumerate(zip(systems, rhs)):
        A64 = I.CSR(Ai.n_rows, Ai.n_cols, Ai.row_ptr, Ai.cols, Ai.vals.astype(np.float64))
        r = bi.astype(np.float64) - O.csr_spmv(A64, xb[offs[i]:offs[i + 1]])
        worst_bd = max(worst_bd, float(np.linalg.norm(r) / np.linalg.norm(bi)))
    out["cfg4b_batch_512"] = {"seconds": float(np.min(ts)), "rows": N,
                              "rows_padded": int(bt.rows_padded), "matrix_bytes": int(bt.matrix_bytes),
                              "nnz": int(sum(a.nnz for a in systems)), "iterations_max": int(it.max()),
                              "iterations_mean": float(it.mean()), "oracle_relres_max": worst,
                              "block_diagonal_single_solver_seconds": float(np.min(tb)),
                              "block_diagonal_iterations": int(stb.iterations),
                              "block_diagonal_converged": bool(stb.converged),
                              "block_diagonal_oracle_relres_max_per_system": worst_bd,
                              "speedup_vs_block_diagonal": float(np.min(tb) / np.min(ts))}
    log(f"cfg4b: {np.min(ts) * 1e3:.1f} ms batched vs {np.min(tb) * 1e3:.1f} ms block-diagonal single solve "
        f"({stb.iterations} it, worst system relres {worst_bd:.2e})")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--droptol", type=float, default=0.125)   # fastainv_sym default (fastainv.h:14)
    ap.add_argument("--spmv-only", action="store_true")
    ap.add_argument("--batching", action="store_true",
                    help="time the 128^3 finite-volume, block-RHS and 512-system batched configurations "
                         "instead (one JSON line of its own)")
    ap.add_argument("--foam", action="store_true",
                    help="time the foam-extend call pattern (lduMatrix input, device coefficient refresh, "
                         "normalised-residual solves, drop tolerance tuned across time steps) on a 128^3 mesh")
    ap.add_argument("--droptol-sweep", action="store_true",
                    help="also time the FP64 and mixed PCG solves with the factors of drop tolerances "
                         "2^-4 and 2^-5 (set-up on the host + solve, one line of its own)")
    ap.add_argument("--spmm", action="store_true",
                    help="multi-vector products of 7- and 27-point 128^3 stencils (pattern and symmetric plans)")
    ap.add_argument("--hblock", action="store_true",
                    help="time the SpMV of a banded two-unknowns-per-node matrix stored as hblock pairs, "
                         "LDS x windows against global gathers (one JSON line of its own)")
    ap.add_argument("--calibrate", action="store_true",
                    help="first copy 1 GiB device-to-device 3 times (a plain copy of known size, "
                         "the reference for rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the library has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)
    import ispm_amd as I
    import ispm_amd.gen as G
    import oracle as O
    I.init()
    log("device", I.device_string())
    if args.calibrate:
        nb = 1 << 30
        src, dst = I.DeviceArray(nb // 8), I.DeviceArray(nb // 8)
        for _ in range(3):
            I.check(I.lib().ispm_memcpy(dst.vp, src.vp, nb, 2))   # ISPM_D2D
        I.synchronize()
        del src, dst
        log("calibration: 3 device-to-device copies of 1 GiB")
    if args.foam:
        print(json.dumps({"metric": "foam-extend call pattern", "legs": foam_legs(I, G, O)}))
        return
    if args.spmm:
        print(json.dumps({"metric": "SpMM of stencil plans, k = 3 and 8",
                          "legs": spmm_legs(I, G, O, max(args.steps, 10), args.warmup)}))
        return
    if args.hblock:
        print(json.dumps({"metric": "hblock SpMV, LDS x windows vs global gathers",
                          "legs": hblock_legs(I, G, O, max(args.steps, 10), args.warmup)}))
        return
    if args.batching:
        from ispm_amd import solver as S
        print(json.dumps({"metric": "batching configurations", "legs": batching_legs(I, G, O, S)}))
        return
    t = time.perf_counter()
    A = G.stencil27_contrast(192)
    log(f"generated 27-point 192^3: n={A.n_rows} nnz={A.nnz} in {time.perf_counter() - t:.1f}s")
    legs = {}
    # fp32_sym: x with guard zones, as the solver holds it (the four-rows-per-lane kernel);
    # fp32_sym_checked: the same plan through the bounds-checked kernel (x without guards)
    for name, vt, et, fl, gd in (("fp64", I.F64, I.F64, 0, False), ("fp32", I.F32, I.F32, 0, False),
                                 ("fp64_sym", I.F64, I.F64, I.SPMV_PLAN_SYMMETRIC, False),
                                 ("fp32_sym", I.F32, I.F32, I.SPMV_PLAN_SYMMETRIC, True),
                                 ("fp32_sym_checked", I.F32, I.F32, I.SPMV_PLAN_SYMMETRIC, False)):
        t = time.perf_counter()
        legs[name] = spmv_leg(I, G, A, vt, et, max(args.steps, 10), args.warmup, fl, guarded=gd)
        log(f"spmv {name}: {legs[name]['GBps']:.0f} GB/s ({legs[name]['avg_s'] * 1e6:.1f} us) "
            f"[{time.perf_counter() - t:.1f}s]")
    if args.spmv_only:
        dom = legs["fp64"]
        print(json.dumps({"metric": "SpMV HBM GB/s", "value": dom["GBps"], "unit": "GB/s",
                          "spmv": legs}))
        return
    from ispm_amd import solver as S
    tol = 1e-10
    t = time.perf_counter()
    factors = I.fastainv_sym(A, I.F64, max_elts=64 * A.n_rows, droptol=args.droptol)
    setup_s = time.perf_counter() - t
    log(f"fastainv_sym droptol={args.droptol}: nnz(Ainv1)/n = {factors[0].nnz / A.n_rows:.2f} in {setup_s:.1f}s")
    b = host_spmv(A, G.uniform(1, A.n_rows))
    solves = {}
    steps, warm = args.steps, args.warmup
    for name, prec, meth in (("fp64_ppcg", "fp64", S.PPCG), ("fp64_pcg", "fp64", S.PCG),
                             ("fp64_pcg1", "fp64", S.PCG1), ("mixed_pcg", "mixed", S.PCG),
                             ("mixed_pcg1", "mixed", S.PCG1), ("mixed_ppcg", "mixed", S.PPCG)):
        t = time.perf_counter()
        solves[name], _ = solve_leg(I, S, O, A, b, factors, prec, meth, steps, warm, tol)
        v = solves[name]
        log(f"solve {name}: {v['seconds'] * 1e3:.1f} ms, {v['iterations']} it, "
            f"{v['replacements']} repl, oracle relres {v['oracle_relres']:.2e} [{time.perf_counter() - t:.1f}s]")
    if args.droptol_sweep:
        # set-up (host fastainv_sym) + solve at smaller drop tolerances: denser factors, fewer
        # iterations, more off-diagonal preconditioner work per iteration
        sweep = {"2^-3": {"setup_s": setup_s, "fsai_nnz_per_row": factors[0].nnz / A.n_rows,
                          "fp64_pcg": solves["fp64_pcg"], "mixed_pcg": solves["mixed_pcg"]}}
        for lg in (-4, -5):
            t = time.perf_counter()
            fac = I.fastainv_sym(A, I.F64, max_elts=64 * A.n_rows, droptol=2.0 ** lg)
            ent = {"setup_s": time.perf_counter() - t, "fsai_nnz_per_row": fac[0].nnz / A.n_rows}
            for name, prec in (("fp64_pcg", "fp64"), ("mixed_pcg", "mixed")):
                ent[name], _ = solve_leg(I, S, O, A, b, fac, prec, S.PCG, min(steps, 3), 1, tol)
            sweep[f"2^{lg}"] = ent
            log(f"droptol 2^{lg}: set-up {ent['setup_s']:.1f} s, nnz(Ainv1)/n {ent['fsai_nnz_per_row']:.2f}; "
                f"fp64 {ent['fp64_pcg']['seconds'] * 1e3:.1f} ms / {ent['fp64_pcg']['iterations']} it, "
                f"mixed {ent['mixed_pcg']['seconds'] * 1e3:.1f} ms / {ent['mixed_pcg']['iterations']} it")
            del fac
        print(json.dumps({"metric": "FSAI-PCG over drop tolerances", "sweep": sweep}))
        return
    # the fastest solve of each precision: the speed-up compares best with best
    fp64_best = min(v["seconds"] for k, v in solves.items() if k.startswith("fp64"))
    mixed = min((v for k, v in solves.items() if k.startswith("mixed")), key=lambda v: v["seconds"])
    cpu = None
    if not args.no_cpu:
        cpu = cpu_solve_baseline(O, A, b, tol, droptol=args.droptol)
        cpu["spmv"] = cpu_spmv_baseline(O, A, seconds=5.0)
        log(f"cpu: solve {cpu['value']:.2f} s ({cpu['iterations']} it, {cpu['cores']} threads), "
            f"fastainv {cpu['fastainv_seconds_1core']:.2f} s")
    out = {
        "metric": "FSAI-PCG time to 1e-10 residual, mixed vs FP64 on 1 GPU; SpMV HBM GB/s",
        "value": mixed["seconds"],
        "unit": "s (mixed-precision FSAI-PCG solve to 1e-10 FP64 relative residual)",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": mixed["seconds"] * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32/bf16 inner, f64 residual",
        "data": "synthetic (seeded 27-point 192^3 stencil, contrast 1e4 in 8^3 blocks; b = A x_true, x_true ~ U[-1,1])",
        "config": {"workload": "27-point 192^3 contrast-1e4, FSAI-PCG to 1e-10", "n_rows": A.n_rows,
                   "nnz": A.nnz, "droptol": args.droptol, "tol": tol},
        "speedup_mixed_vs_fp64": fp64_best / mixed["seconds"],
        "fsai_setup_s": setup_s,
        "fsai_nnz_per_row": factors[0].nnz / A.n_rows,
        "solves": solves,
        "spmv": legs,
        # the dominant kernel of the timed (mixed) solve: its A product
        "roofline": (roofline(legs["fp32_sym"], "fp32_sym",
                              "symdia_r4_kernel<1,14> (FP32 symmetric-storage 27-point 192^3 "
                              "SpMV on x with guard zones, four rows per lane; the largest "
                              "share of the mixed solve)")
                     if mixed["A_storage"] != "SlELL" else
                     roofline(legs["fp32"], "fp32",
                              "slell_diagpat_kernel<float,float> (FP32 27-point 192^3 SpMV, the "
                              "largest share of the mixed solve)")),
        "roofline_fp64_spmv": roofline(legs["fp64"], "fp64",
                                       "slell_diagpat_kernel<double,double> (FP64 27-point 192^3 SpMV)"),
        "roofline_fp64_sym_spmv": roofline(legs["fp64_sym"], "fp64_sym",
                                           "symdia_kernel<double,double> (FP64 symmetric-storage "
                                           "27-point 192^3 SpMV, the A product of the FP64 solves)"),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
