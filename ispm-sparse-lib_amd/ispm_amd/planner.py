"""Learned SpMV-plan and FSAI drop-tolerance predictor.

A 32-feature -> 64 -> 64 -> (C plan scores, 1 log2-droptol) MLP whose forward, backward and
Adam kernels are the library's own (csrc/mlp.hip, ispm_mlp_*) reads cheap statistics of a CSR
matrix and predicts
  * which SlELL format (slice height S, variable height V, diagonal slices D, hblock W) runs
    the SpMV fastest: a distribution over CLASSES, trained towards a soft target built from
    the measured SpMV time of every class (the exhaustive autotune's table,
    ispm_plan_time_formats), and
  * the FSAI drop tolerance that minimises factor set-up plus FSAI-PCG solve time, the
    quantity the reference's DroptolTuner (fastainv/droptol-tuning.h) searches for by golden
    section over log2 droptol, trained on measured times over DROPTOL_GRID.
Decisions fall back to the exhaustive autotune when the plan prediction is not confident and
to the default drop tolerance 0.125 (fastainv/fastainv.h:14 of the reference) when the
droptol prediction leaves the trained range; the same two procedures produce the training
labels and judge the model on held-out matrices (evaluate()).

Command line (each step resumable; corpus and model files are small .npz):
    python -m ispm_amd.planner label --first 0 --count 700 --workers 12 --out corpus_a.npz
    python -m ispm_amd.planner train --corpus corpus_a.npz corpus_b.npz --holdout 200 --out model.npz
    python -m ispm_amd.planner eval --corpus ... --model model.npz --holdout 200
Labelling runs `--workers` processes that share the one GPU: matrix generation, features,
layout conversions and the host factorisations run in parallel, while every upload, timed
launch and timed solve holds an advisory file lock, so no two timings overlap on the device.
"""
from __future__ import annotations

import argparse
import contextlib
import ctypes as C
import fcntl
import functools
import os
import sys
import tempfile
import time

import numpy as np

from . import (CSR, F64, DeviceArray, check, fastainv_sym, lib, synchronize)
from . import gen as G

N_FEATURES = 32
HIDDEN = 64
DEFAULT_DROPTOL = 0.125
# log2 drop tolerances timed for labels: the DroptolTuner's bracket [-10, -3]
# (fastainv/droptol-tuning.h:23), largest first; grid[0] is the default 0.125
DROPTOL_GRID = tuple(float(-k) for k in range(3, 11))
CORPUS_SIZE = 2000
HOLDOUT = 200


def _classes():
    out = []
    for S in (4, 8, 16, 32, 64, 128, 256, 512):
        for V in ((0,) if S < 32 else (0, 1)):
            for D, W in ((0, 1), (1, 1), (0, 2)):
                out.append((S, V, D, W))
    return out


CLASSES = _classes()
N_CLASSES = len(CLASSES)
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "planner_model.npz")


# --------------------------------------------------------------------------- features --
def _ell_slots(rl, S):
    pad = (-rl.size) % S
    r = np.concatenate([rl, np.zeros(pad, rl.dtype)]).reshape(-1, S)
    return float(r.max(1).sum() * S)


def features(A: CSR) -> np.ndarray:
    """32 statistics of a CSR matrix, O(nnz) numpy: sizes, row-length moments and histogram,
    bandwidth, diagonal structure and pairs (hblock potential), diagonal dominance,
    coefficient contrast, and the plain-ELL padding at five slice heights."""
    n, nnz = A.n_rows, A.nnz
    rl = np.diff(A.row_ptr.astype(np.int64))
    rows = np.repeat(np.arange(n, dtype=np.int64), rl)
    cols = A.cols.astype(np.int64)
    vals = np.abs(A.vals.astype(np.float64))
    f = np.zeros(N_FEATURES)
    f[0] = np.log2(max(n, 1)) / 24
    f[1] = np.log2(max(nnz, 1)) / 30
    f[2] = rl.mean() / 32
    f[3] = rl.std() / 32
    f[4] = np.log2(1 + rl.max()) / 10
    f[5] = np.log2(1 + rl.min()) / 10
    h, _ = np.histogram(rl, bins=[0, 1, 2, 4, 8, 16, 32, 64, 1 << 30])
    f[6:14] = h / max(n, 1)
    off = cols - rows
    dist = np.abs(off)
    if nnz:
        f[14] = np.log2(1 + dist.mean()) / 24
        f[15] = np.log2(1 + dist.max()) / 24
        f[16] = np.mean(dist <= 1)
        u, c = np.unique(off, return_counts=True)
        f[17] = np.log2(1 + u.size) / 24
        f[18] = np.sort(c)[::-1][:27].sum() / nnz          # share of the 27 densest diagonals
    if nnz > 1:
        same = rows[1:] == rows[:-1]
        f[19] = np.mean(same & (cols[:-1] % 2 == 0) & (cols[1:] == cols[:-1] + 1))
    isd = cols == rows
    d = np.zeros(n)
    np.add.at(d, rows[isd], vals[isd])
    offsum = np.zeros(n)
    np.add.at(offsum, rows[~isd], vals[~isd])
    dom = d / np.maximum(offsum, 1e-300)
    f[20] = np.clip(np.log10(np.median(dom) + 1e-12), -3, 3) / 3
    f[21] = np.clip(np.log10(dom.min() + 1e-12), -3, 3) / 3
    nzv = vals[vals > 0]
    f[22] = np.log10(nzv.max() / nzv.min()) / 12 if nzv.size else 0
    dp = d[d > 0]
    f[23] = np.log10(dp.max() / dp.min()) / 12 if dp.size else 0
    ratio = vals[~isd] / np.maximum(d[rows[~isd]], 1e-300)
    if ratio.size:
        f[24] = np.clip(np.log10(np.median(ratio) + 1e-12), -6, 1) / 6
        f[25] = np.clip(np.log10(np.percentile(ratio, 90) + 1e-12), -6, 1) / 6
    for j, S in enumerate((4, 16, 64, 256, 512)):
        f[26 + j] = np.log2(_ell_slots(rl, S) / max(nnz, 1)) / 4 if nnz else 0
    f[31] = 1.0
    return f.astype(np.float32)


# ----------------------------------------------------------------------- GPU labels --
def time_formats(A: CSR, vt=F64, et=F64, threads=0, lock_fd=-1) -> np.ndarray:
    """Measured SpMV seconds of every class (default launch of each format; +inf where a
    format cannot be built): the exhaustive autotune's table over CLASSES.  lock_fd >= 0: an
    open file whose flock() lock is held while layouts are uploaded and timed (processes
    sharing the GPU); threads: host conversion threads (0: up to 8)."""
    S = np.array([c[0] for c in CLASSES], np.uint32)
    V = np.array([c[1] for c in CLASSES], np.int32)
    D = np.array([c[2] for c in CLASSES], np.int32)
    W = np.array([c[3] for c in CLASSES], np.uint32)
    out = np.full(N_CLASSES, np.inf)
    check(lib().ispm_plan_time_formats_shared(vt, et, C.byref(A.c()), N_CLASSES, S.ctypes.data,
                                              W.ctypes.data, V.ctypes.data, D.ctypes.data,
                                              out.ctypes.data, int(threads), int(lock_fd)))
    return out


@contextlib.contextmanager
def device_lock(fd):
    """Exclusive use of the GPU among the processes locking the same file (fd < 0: none)."""
    if fd is None or fd < 0:
        yield
        return
    fcntl.flock(fd, fcntl.LOCK_EX)
    try:
        yield
    finally:
        fcntl.flock(fd, fcntl.LOCK_UN)


def plan_target(times, tau: float = 8.0) -> np.ndarray:
    """Soft target over the classes: softmax of -tau * log(t / t_best)."""
    t = np.asarray(times, np.float64)
    ok = np.isfinite(t)
    s = np.full(t.size, -np.inf)
    s[ok] = -tau * np.log(t[ok] / t[ok].min())
    e = np.where(ok, np.exp(s - s[ok].max()), 0.0)
    return (e / e.sum()).astype(np.float32)


def droptol_times(A: CSR, grid=DROPTOL_GRID, tol: float = 1e-8, seed: int = 77, lock_fd=-1,
                  max_setup: float = 120.0):
    """Seconds of factor set-up (fastainv_sym on the host) plus FP64 FSAI-PCG solve on the GPU
    (timed after one warm-up solve) at each log2 drop tolerance of the grid, largest first.

    Returns (total, setup, iters), one entry per grid point.  Set-up time grows as the drop
    tolerance falls (more fill), so once a factorisation alone takes at least the best total so
    far, no smaller drop tolerance can win: the search stops there (also past `max_setup`
    seconds).  Entries not solved have total = +inf; their `setup` holds the last measured
    set-up (a lower bound of their own) and iters = 0.  A solve that does not converge has
    total = +inf and iters = -1."""
    from .solver import PCG, Solver
    xt = G.uniform(seed, A.n_rows)
    b = np.asarray(A.to_scipy() @ xt)
    total = np.full(len(grid), np.inf)
    setup = np.full(len(grid), np.nan)
    iters = np.zeros(len(grid), np.int32)
    best, A_plan, bound = np.inf, None, np.nan
    for k, lg in enumerate(grid):
        t0 = time.perf_counter()
        M = fastainv_sym(A, F64, max_elts=64 * A.n_rows, droptol=2.0 ** lg)
        setup[k] = bound = time.perf_counter() - t0
        if setup[k] >= best:
            break
        with device_lock(lock_fd):
            sol = Solver(A, M=M, precision="fp64", method=PCG, tol=tol, maxit=20000, fsai_product=False,
                         A_plan=A_plan)
            A_plan = sol.A_plan
            n = sol.n
            bd = DeviceArray.from_numpy(np.concatenate([b, np.zeros(n - A.n_rows)]))
            xd = DeviceArray(n)
            sol.solve_device(bd, xd, allow_fail=True)
            xd.zero()
            synchronize()
            t0 = time.perf_counter()
            st = sol.solve_device(bd, xd, allow_fail=True)
            synchronize()
            ts = time.perf_counter() - t0
            del sol, bd, xd
        if st.converged:
            total[k] = setup[k] + ts
            iters[k] = st.iterations
            best = min(best, total[k])
        else:
            iters[k] = -1
        if setup[k] > max_setup:
            break
    setup[np.isnan(setup)] = bound
    return total, setup, iters


# ----------------------------------------------------------------------------- corpus --
def corpus_spec(i: int) -> dict:
    """Deterministic description of corpus matrix i: rows log-uniform in [1e4, 2e6]; 3 in 4
    are 7/19/27-point stencils on boxes of random aspect (z anisotropy 1..100 and coefficient
    contrast 1..1e6, log-uniform), the rest random geometric kNN graph Laplacians with 10..26
    neighbours."""
    u = G.uniform01(900001 + i, 8)
    n = int(10 ** (4 + u[2] * np.log10(200.0)))
    if u[0] < 0.75:
        pts = (7, 19, 27)[min(2, int(u[1] * 3))]
        nx = max(8, int(round(n ** (1 / 3) * (0.7 + 0.6 * u[3]))))
        ny = max(8, int(round(n ** (1 / 3) * (0.7 + 0.6 * u[4]))))
        nz = max(8, n // (nx * ny))
        return dict(kind="stencil", nx=nx, ny=ny, nz=nz, points=pts, anisotropy=float(10 ** (2 * u[5])),
                    contrast=float(10 ** (6 * u[6])), seed=1000 + i)
    return dict(kind="rgg", n=n, k=10 + min(16, int(u[3] * 17)), seed=2000 + i)


def spec_rows(spec: dict) -> int:
    return spec["nx"] * spec["ny"] * spec["nz"] if spec["kind"] == "stencil" else spec["n"]


def make_matrix(spec: dict) -> CSR:
    if spec["kind"] == "stencil":
        return G.stencil(spec["nx"], spec["ny"], spec["nz"], spec["points"], spec["anisotropy"],
                         spec["contrast"], seed=spec["seed"])
    return G.rgg_laplacian(spec["n"], spec["k"], spec["seed"])


_WORKER = {}


def _worker_init(lock_path, threads):
    from . import init
    init()
    _WORKER["fd"] = os.open(lock_path, os.O_RDWR | os.O_CREAT, 0o600) if lock_path else -1
    _WORKER["threads"] = threads


def label_one(i: int, droptol=True, lock_fd=None, threads=None):
    """(i, features, format times, droptol totals, set-up seconds, iterations, log line) of
    corpus matrix i."""
    lock_fd = _WORKER.get("fd", -1) if lock_fd is None else lock_fd
    threads = _WORKER.get("threads", 0) if threads is None else threads
    t0 = time.perf_counter()
    A = make_matrix(corpus_spec(i))
    t1 = time.perf_counter()
    x = features(A)
    t2 = time.perf_counter()
    T = time_formats(A, threads=threads, lock_fd=lock_fd)
    t3 = time.perf_counter()
    if droptol:
        DT, DS, DI = droptol_times(A, lock_fd=lock_fd)
    else:
        DT = np.full(len(DROPTOL_GRID), np.inf)
        DS = np.full(len(DROPTOL_GRID), np.nan)
        DI = np.zeros(len(DROPTOL_GRID), np.int32)
    dbest = f"2^{DROPTOL_GRID[int(np.argmin(DT))]:.0f}" if np.isfinite(DT).any() else "n/a"
    line = (f"matrix {i}: n={A.n_rows} nnz={A.nnz} best {CLASSES[int(np.argmin(T))]} "
            f"{np.min(T) * 1e6:.1f} us, droptol {dbest} ({int((DI != 0).sum())} solved) "
            f"[{time.perf_counter() - t0:.1f}s: matrix {t1 - t0:.1f}, features {t2 - t1:.1f}, "
            f"formats {t3 - t2:.1f}, droptol {time.perf_counter() - t3:.1f}]")
    return i, x, T, DT, DS, DI, line


def _save(out, rows):
    rows = sorted(rows, key=lambda r: r[0])
    tmp = out + ".tmp.npz"
    np.savez(tmp, X=np.array([r[1] for r in rows], np.float32), T=np.array([r[2] for r in rows]),
             DT=np.array([r[3] for r in rows]), DS=np.array([r[4] for r in rows]),
             DI=np.array([r[5] for r in rows], np.int32), idx=np.array([r[0] for r in rows]),
             droptol_grid=np.array(DROPTOL_GRID))
    os.replace(tmp, out)      # a run stopped mid-write leaves the previous file whole


def label(first: int, count: int, out: str, log=print, droptol=True, workers=1, threads=0,
          deadline=None):
    """Features and measured labels of corpus matrices first .. first+count-1, saved to `out`
    after every matrix (a partial file is resumed).  workers > 1: that many processes share
    the GPU through a file lock (largest matrices first, so the tail is short).  deadline:
    seconds after which no further result is awaited (unfinished matrices stay unlabelled)."""
    rows = []
    if os.path.exists(out):
        z = np.load(out, allow_pickle=False)
        rows = [(int(z["idx"][k]), z["X"][k], z["T"][k], z["DT"][k], z["DS"][k], z["DI"][k])
                for k in range(z["idx"].size)]
    done = set(r[0] for r in rows)
    todo = [i for i in range(first, first + count) if i not in done]
    todo.sort(key=lambda i: -spec_rows(corpus_spec(i)))
    t0 = time.perf_counter()
    if workers <= 1:
        from . import init
        init()
        results = (label_one(i, droptol, -1, threads) for i in todo)
        pool = None
    else:
        import multiprocessing as mp
        lock_path = os.path.join(tempfile.gettempdir(), f"ispm_label_{os.getpid()}.lock")
        pool = mp.get_context("spawn").Pool(workers, initializer=_worker_init,
                                             initargs=(lock_path, threads))
        results = pool.imap_unordered(functools.partial(label_one, droptol=droptol), todo)
    stopped = False
    try:
        for r in results:
            rows.append(r[:6])
            _save(out, rows)
            log(f"{r[6]} ({len(rows)} labelled, {time.perf_counter() - t0:.0f}s)")
            if deadline is not None and time.perf_counter() - t0 > deadline:
                stopped = True
                log(f"deadline reached: {len(rows)} labelled, the rest left for a later run")
                break
    finally:
        if pool is not None:
            if stopped:
                pool.terminate()
            else:
                pool.close()
            pool.join()
    return out


def load_corpus(paths):
    """(X, T, DT, DS, DI, idx) of the corpus files, ordered by matrix index."""
    parts = {k: [] for k in ("X", "T", "DT", "DS", "DI", "idx")}
    for p in paths:
        z = np.load(p, allow_pickle=False)
        if "droptol_grid" not in z.files or not np.array_equal(z["droptol_grid"], np.array(DROPTOL_GRID)):
            raise ValueError(f"{p}: labelled over another drop-tolerance grid")
        for k in parts:
            parts[k].append(z[k])
    idx = np.concatenate(parts["idx"])
    order = np.argsort(idx)
    return tuple(np.concatenate(parts[k])[order] for k in ("X", "T", "DT", "DS", "DI")) + (idx[order],)


def targets(T, DT, tau: float = 8.0):
    """(Tp, Mp, Td, Md) training targets from the timing tables."""
    N = T.shape[0]
    Tp = np.zeros((N, N_CLASSES), np.float32)
    Mp = np.zeros(N, np.float32)
    Td = np.zeros(N, np.float32)
    Md = np.zeros(N, np.float32)
    for i in range(N):
        if np.isfinite(T[i]).any():
            Tp[i] = plan_target(T[i], tau)
            Mp[i] = 1
        ok = np.isfinite(DT[i])
        if ok.any():
            Td[i] = DROPTOL_GRID[int(np.argmin(np.where(ok, DT[i], np.inf)))]
            Md[i] = 1
    return Tp, Mp, Td, Md


# ---------------------------------------------------------------------------- model --
def _dp(a: DeviceArray, off_elems: int = 0):
    return C.c_void_p(a.ptr + off_elems * a.dtype.itemsize)


class Model:
    """Parameters (flat FP32, the layout of ispm_mlp_*) and the feature normalisation."""

    def __init__(self, params=None, mean=None, std=None, seed=0, confidence=0.3, droptol_offset=0.0):
        self.confidence = float(confidence)
        self.droptol_offset = float(droptol_offset)
        self.n_params = lib().ispm_mlp_param_count(N_CLASSES)
        if params is None:
            rng = np.random.default_rng(seed)
            parts = []
            for fan_in, fan_out in ((N_FEATURES, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, N_CLASSES + 1)):
                parts.append(rng.normal(0, np.sqrt(2.0 / fan_in), fan_out * fan_in))
                parts.append(np.zeros(fan_out))
            params = np.concatenate(parts)
        self.params = np.asarray(params, np.float32)
        if self.params.size != self.n_params:
            raise ValueError("parameter count does not match the network")
        self.mean = np.zeros(N_FEATURES, np.float32) if mean is None else np.asarray(mean, np.float32)
        self.std = np.ones(N_FEATURES, np.float32) if std is None else np.asarray(std, np.float32)

    def normalize(self, X):
        return ((np.atleast_2d(np.asarray(X, np.float32)) - self.mean) / self.std).astype(np.float32)

    def forward(self, X):
        """Raw outputs (B x (C+1)) from the GPU forward kernels."""
        Xn = self.normalize(X)
        B = Xn.shape[0]
        p = DeviceArray.from_numpy(self.params)
        xd = DeviceArray.from_numpy(Xn.ravel())
        out = DeviceArray(B * (N_CLASSES + 1), np.float32)
        check(lib().ispm_mlp_forward(N_CLASSES, p.vp, xd.vp, B, out.vp))
        return out.numpy().reshape(B, N_CLASSES + 1)

    def predict(self, X):
        """(class probabilities B x C, log2 droptol B: the regression output plus the
        calibrated offset)."""
        z = self.forward(X)
        zc = z[:, :N_CLASSES].astype(np.float64)
        p = np.exp(zc - zc.max(1, keepdims=True))
        p /= p.sum(1, keepdims=True)
        return p, z[:, N_CLASSES] + self.droptol_offset

    def fit(self, X, Tp, Mp, Td, Md, epochs=400, batch=64, lr=3e-3, drop_weight=0.5, seed=0, log=None):
        """Adam on the GPU (ispm_mlp_train_step); the shuffled set is uploaded once per epoch
        and batches are contiguous slices of it.  Returns the loss of every epoch's last batch."""
        X = np.asarray(X, np.float32)
        self.mean = X.mean(0)
        self.std = X.std(0) + 1e-3
        Xn = self.normalize(X)
        N = Xn.shape[0]
        B = min(batch, N)
        rng = np.random.default_rng(seed)
        p = DeviceArray.from_numpy(self.params)
        m = DeviceArray(self.n_params, np.float32)
        v = DeviceArray(self.n_params, np.float32)
        loss = DeviceArray(1, np.float32)
        dev = [DeviceArray(N * w, np.float32) for w in (N_FEATURES, N_CLASSES, 1, 1, 1)]
        hist, step = [], 0
        for ep in range(epochs):
            order = rng.permutation(N)
            for d, a in zip(dev, (Xn, Tp, Mp, Td, Md)):
                d.upload(np.ascontiguousarray(np.asarray(a, np.float32)[order]).ravel())
            for b0 in range(0, N - B + 1, B):
                step += 1
                check(lib().ispm_mlp_train_step(
                    N_CLASSES, p.vp, m.vp, v.vp, step, _dp(dev[0], b0 * N_FEATURES), _dp(dev[1], b0 * N_CLASSES),
                    _dp(dev[2], b0), _dp(dev[3], b0), _dp(dev[4], b0), B, C.c_float(lr), C.c_float(drop_weight),
                    loss.vp))
            hist.append(float(loss.numpy()[0]))
            if log and (ep % 50 == 0 or ep == epochs - 1):
                log(f"epoch {ep}: loss {hist[-1]:.4f}")
        self.params = p.numpy()
        return hist

    def calibrate(self, X, quantile=0.25):
        """Confidence threshold: the given quantile of the top-class probability over X (the
        share of matrices below it goes to the exhaustive fallback)."""
        p, _ = self.predict(X)
        self.confidence = float(np.quantile(p.max(1), quantile))
        return self.confidence

    def calibrate_droptol(self, X, DT, DS, DI, offsets=np.arange(-1.0, 1.0001, 0.05)):
        """Offset added to the regression output before it is rounded to a grid point: the
        one (smallest in size among equals) that minimises the geometric mean of set-up +
        solve time over the best grid point on X.  The regression is trained towards grid
        labels that are mostly the default, so its output for the rest lies between grid
        points; the offset places the rounding threshold where the measured cost says."""
        self.droptol_offset = 0.0
        _, lg = self.predict(X)
        grid = np.array(DROPTOL_GRID)
        cost = droptol_cost(DT, DS, DI)
        best = np.min(np.where(np.isfinite(DT), DT, np.inf), 1)
        ok = np.isfinite(best)
        rows = np.arange(len(lg))
        res = []
        for off in offsets:
            k = np.abs(np.clip(lg + off, grid.min(), grid.max())[:, None] - grid[None, :]).argmin(1)
            r = np.minimum(cost[rows, k] / best, 1e3)
            res.append((float(np.exp(np.mean(np.log(r[ok])))), abs(float(off)), float(off)))
        self.droptol_offset = min(res)[2]
        return self.droptol_offset

    def save(self, path):
        np.savez(path, params=self.params, mean=self.mean, std=self.std, confidence=self.confidence,
                 droptol_offset=self.droptol_offset, classes=np.array(CLASSES, np.int32),
                 droptol_grid=np.array(DROPTOL_GRID))

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        if not np.array_equal(z["classes"], np.array(CLASSES, np.int32)):
            raise ValueError("the model was trained on another class list")
        if "droptol_grid" in z.files and not np.array_equal(z["droptol_grid"], np.array(DROPTOL_GRID)):
            raise ValueError("the model was trained on another drop-tolerance grid")
        conf = float(z["confidence"]) if "confidence" in z.files else 0.3
        off = float(z["droptol_offset"]) if "droptol_offset" in z.files else 0.0
        return cls(z["params"], z["mean"], z["std"], confidence=conf, droptol_offset=off)


def load_default():
    return Model.load(MODEL_PATH) if os.path.exists(MODEL_PATH) else None


# ---------------------------------------------------------------------- decisions --
def choose_plan(A: CSR, vt=F64, et=F64, model=None, confidence=None):
    """(S, V, D, hblock, source): the model's class when its probability reaches
    `confidence` (default: the model's calibrated threshold), else the fastest class of the
    exhaustive timing table."""
    model = model or load_default()
    if model is not None:
        p, _ = model.predict(features(A))
        c = int(np.argmax(p[0]))
        if p[0, c] >= (model.confidence if confidence is None else confidence):
            return (*CLASSES[c], "model")
    t = time_formats(A, vt, et)
    return (*CLASSES[int(np.argmin(t))], "exhaustive")


def choose_droptol(A: CSR, model=None):
    """(droptol, source): the grid point nearest the model's prediction when that lies inside
    the trained range (half a step of slack), else the default 0.125."""
    model = model or load_default()
    if model is not None:
        _, lg = model.predict(features(A))
        lg = float(lg[0])
        grid = np.array(DROPTOL_GRID)
        if grid.min() - 0.5 <= lg <= grid.max() + 0.5:
            return 2.0 ** float(grid[np.abs(grid - lg).argmin()]), "model"
    return DEFAULT_DROPTOL, "default"


def best_fixed_class(T) -> int:
    """The one class with the smallest geometric-mean time ratio to the best over a set: the
    non-learned baseline a single hand-picked format would give."""
    return int(np.argmin(np.exp(np.log(T / T.min(1, keepdims=True)).mean(0))))


def droptol_cost(DT, DS, DI):
    """Per grid point: the measured total where solved; the set-up lower bound where the
    search stopped before solving (the total there is at least that, and above the best);
    +inf where the solve did not converge."""
    return np.where(np.isfinite(DT), DT, np.where(DI == 0, DS, np.inf))


def evaluate(model, X, T, DT=None, DS=None, DI=None, confidence=None, fixed_class=None):
    """Held-out quality: for the plan head, the time of the chosen class (model when
    confident, else exhaustive) over the exhaustive best, and of the model's top class alone
    (and of one fixed class, when given); for the droptol head, set-up + solve time at the grid
    point nearest the prediction over that at the default 0.125 and over the best grid point,
    over all matrices and over the slice whose best grid point is not the default (a
    prediction on a point the label search skipped counts its set-up lower bound)."""
    p, lg = model.predict(X)
    confidence = model.confidence if confidence is None else confidence
    res = {"n": int(X.shape[0]), "confidence_threshold": confidence}
    if fixed_class is not None:
        r = T[:, fixed_class] / T.min(1)
        res["plan_fixed_class"] = CLASSES[fixed_class]
        res["plan_fixed_ratio_gmean"] = float(np.exp(np.mean(np.log(r))))
        res["plan_fixed_within_5pct"] = float(np.mean(r <= 1.05))
    top = p.argmax(1)
    best = T.min(1)
    t_top = T[np.arange(len(top)), top]
    conf = p.max(1) >= confidence
    t_choice = np.where(conf, t_top, best)
    res["plan_top_ratio_gmean"] = float(np.exp(np.mean(np.log(t_top / best))))
    res["plan_top_within_5pct"] = float(np.mean(t_top <= 1.05 * best))
    res["plan_choice_ratio_gmean"] = float(np.exp(np.mean(np.log(t_choice / best))))
    res["plan_choice_within_5pct"] = float(np.mean(t_choice <= 1.05 * best))
    res["plan_confident_fraction"] = float(np.mean(conf))
    if DT is None:
        return res
    if DS is None:
        DS = np.where(np.isfinite(DT), 0.0, np.inf)
    if DI is None:
        DI = np.where(np.isfinite(DT), 1, -1)
    grid = np.array(DROPTOL_GRID)
    cost = droptol_cost(DT, DS, DI)
    k = np.abs(np.clip(lg, grid.min(), grid.max())[:, None] - grid[None, :]).argmin(1)
    rows = np.arange(len(k))
    t_pred = cost[rows, k]
    t_def = cost[:, 0]                 # grid[0] = log2(0.125), the default
    t_best = np.min(np.where(np.isfinite(DT), DT, np.inf), 1)
    ok = np.isfinite(t_def) & np.isfinite(t_best) & np.isfinite(t_pred)
    res["droptol_n"] = int(ok.sum())
    res["droptol_pred_not_converged"] = int((np.isfinite(t_best) & ~np.isfinite(t_pred)).sum())
    res["droptol_pred_exact_grid_point"] = float(np.mean(k == np.argmin(np.where(np.isfinite(DT), DT, np.inf), 1)))
    if ok.any():
        gm = lambda a: float(np.exp(np.mean(np.log(a))))
        res["droptol_pred_over_default_gmean"] = gm(t_pred[ok] / t_def[ok])
        res["droptol_pred_over_best_gmean"] = gm(t_pred[ok] / t_best[ok])
        res["droptol_default_over_best_gmean"] = gm(t_def[ok] / t_best[ok])
        sl = ok & (np.argmin(np.where(np.isfinite(DT), DT, np.inf), 1) != 0)
        res["droptol_slice_n"] = int(sl.sum())
        if sl.any():
            res["droptol_slice_pred_over_default_gmean"] = gm(t_pred[sl] / t_def[sl])
            res["droptol_slice_default_over_best_gmean"] = gm(t_def[sl] / t_best[sl])
            res["droptol_slice_pred_over_default_max"] = float(np.max(t_pred[sl] / t_def[sl]))
    return res


# --------------------------------------------------------------------------- command --
def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m ispm_amd.planner")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("label")
    a.add_argument("--first", type=int, default=0)
    a.add_argument("--count", type=int, default=100)
    a.add_argument("--out", required=True)
    a.add_argument("--no-droptol", action="store_true")
    a.add_argument("--workers", type=int, default=1)
    a.add_argument("--threads", type=int, default=0, help="host conversion threads per process")
    a.add_argument("--deadline", type=float, default=None, help="seconds; stop awaiting results after")
    t = sub.add_parser("train")
    t.add_argument("--corpus", nargs="+", required=True)
    t.add_argument("--holdout", type=int, default=HOLDOUT)
    t.add_argument("--epochs", type=int, default=400)
    t.add_argument("--tau", type=float, default=8.0)
    t.add_argument("--lr", type=float, default=3e-3)
    t.add_argument("--out", default=MODEL_PATH)
    e = sub.add_parser("eval")
    e.add_argument("--corpus", nargs="+", required=True)
    e.add_argument("--holdout", type=int, default=HOLDOUT)
    e.add_argument("--model", default=MODEL_PATH)
    args = ap.parse_args(argv)

    def log(*x):
        print(*x, flush=True)
    if args.cmd == "label":
        # the processes that time on the GPU initialise it themselves; this one stays off it
        label(args.first, args.count, args.out, log=log, droptol=not args.no_droptol,
              workers=args.workers, threads=args.threads, deadline=args.deadline)
        return 0
    from . import init
    init()
    X, T, DT, DS, DI, idx = load_corpus(args.corpus)
    ntr = X.shape[0] - args.holdout
    if args.cmd == "train":
        Tp, Mp, Td, Md = targets(T, DT, args.tau)
        model = Model(seed=0)
        model.fit(X[:ntr], Tp[:ntr], Mp[:ntr], Td[:ntr], Md[:ntr], epochs=args.epochs, lr=args.lr, log=log)
        model.calibrate(X[:ntr])
        log("droptol offset", model.calibrate_droptol(X[:ntr], DT[:ntr], DS[:ntr], DI[:ntr]))
        model.save(args.out)
        fixed = best_fixed_class(T[:ntr])
        log("train", evaluate(model, X[:ntr], T[:ntr], DT[:ntr], DS[:ntr], DI[:ntr], fixed_class=fixed))
        log("holdout", evaluate(model, X[ntr:], T[ntr:], DT[ntr:], DS[ntr:], DI[ntr:], fixed_class=fixed))
        return 0
    model = Model.load(args.model)
    log("holdout", evaluate(model, X[ntr:], T[ntr:], DT[ntr:], DS[ntr:], DI[ntr:],
                            fixed_class=best_fixed_class(T[:ntr])))
    return 0


if __name__ == "__main__":
    sys.exit(main())
