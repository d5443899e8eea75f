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


def time_spmv(I, plan, x, y, steps, warmup):
    """Average launch duration of the plan's SpMV kernel, timed with HIP events recorded on
    the library stream (the stream the kernel runs on)."""
    for _ in range(warmup):
        plan.execute(x, y)
    I.synchronize()
    e0, e1 = I.Event(), I.Event()
    ts = []
    for _ in range(steps):
        e0.record()
        plan.execute(x, y)
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_ms(e1) * 1e-3)
    return float(np.mean(ts)), float(np.median(ts))


def spmv_leg(I, G, A, vt, et, steps, warmup):
    plan = I.Plan(A, vt, elm_type=et)
    inf = plan.info()
    n = plan.n_rows_padded
    dt = np.float64 if vt == I.F64 else np.float32
    x = I.DeviceArray.from_numpy(np.concatenate([G.uniform(0, A.n_rows), np.zeros(n - A.n_rows)]).astype(dt))
    y = I.DeviceArray(n, dt)
    mean, med = time_spmv(I, plan, x, y, steps, warmup)
    stored = inf.stored_bytes
    return {
        "format": dict(S=inf.S, hblock=inf.hblock, V=inf.V, D=inf.D,
                       N=inf.launch.N, H=inf.launch.H),
        "avg_s": mean, "median_s": med,
        "stored_bytes": stored, "csr_bytes": inf.csr_bytes,
        "GBps": stored / mean * 1e-9, "csr_equiv_GBps": inf.csr_bytes / mean * 1e-9,
        "GFLOPs": 2.0 * A.nnz / mean * 1e-9,
        "frac_hbm": stored / mean * 1e-9 / HBM_PEAK_GBS,
    }


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
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
    t = time.perf_counter()
    A = G.stencil27_contrast(192)
    log(f"generated 27-point 192^3: n={A.n_rows} nnz={A.nnz} in {time.perf_counter() - t:.1f}s")
    legs = {}
    for name, vt, et in (("fp64", I.F64, I.F64), ("fp32", I.F32, I.F32)):
        t = time.perf_counter()
        legs[name] = spmv_leg(I, G, A, vt, et, max(args.steps, 10), args.warmup)
        log(f"spmv {name}: {legs[name]['GBps']:.0f} GB/s ({legs[name]['avg_s'] * 1e6:.1f} us) "
            f"[{time.perf_counter() - t:.1f}s]")
    dom = legs["fp64"]
    cpu = None if args.no_cpu else cpu_spmv_baseline(O, A)
    out = {
        "metric": "FSAI-PCG time to 1e-10 residual, mixed vs FP64 on 1 GPU; SpMV HBM GB/s",
        "value": dom["GBps"],
        "unit": "GB/s (SpMV, stored-format bytes)",
        "n_gpus": 1,
        "steps": max(args.steps, 10),
        "warmup": args.warmup,
        "ms_per_step": dom["avg_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 27-point 192^3 stencil, contrast 1e4 in 8^3 blocks)",
        "config": {"workload": "27-point 192^3 contrast-1e4 SpMV", "n_rows": A.n_rows, "nnz": A.nnz},
        "spmv": legs,
        "roofline": {"bound": "hbm", "achieved": dom["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["frac_hbm"], "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
