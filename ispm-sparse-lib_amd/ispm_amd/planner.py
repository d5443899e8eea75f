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
    python -m ispm_amd.planner label --first 0 --count 250 --out corpus_a.npz
    python -m ispm_amd.planner train --corpus corpus_a.npz corpus_b.npz --holdout 60 --out model.npz
    python -m ispm_amd.planner eval --corpus ... --model model.npz --holdout 60
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import sys
import time

import numpy as np

from . import (CSR, F64, DeviceArray, check, fastainv_sym, lib, synchronize)
from . import gen as G

N_FEATURES = 32
HIDDEN = 64
DEFAULT_DROPTOL = 0.125
DROPTOL_GRID = (-3.0, -4.0, -5.0, -6.0, -7.0)    # log2 drop tolerances timed for labels


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
def time_formats(A: CSR, vt=F64, et=F64) -> np.ndarray:
    """Measured SpMV seconds of every class (default launch of each format; +inf where a
    format cannot be built): the exhaustive autotune's table over CLASSES."""
    S = np.array([c[0] for c in CLASSES], np.uint32)
    V = np.array([c[1] for c in CLASSES], np.int32)
    D = np.array([c[2] for c in CLASSES], np.int32)
    W = np.array([c[3] for c in CLASSES], np.uint32)
    out = np.full(N_CLASSES, np.inf)
    check(lib().ispm_plan_time_formats(vt, et, C.byref(A.c()), N_CLASSES, S.ctypes.data, W.ctypes.data,
                                       V.ctypes.data, D.ctypes.data, out.ctypes.data))
    return out


def plan_target(times, tau: float = 8.0) -> np.ndarray:
    """Soft target over the classes: softmax of -tau * log(t / t_best)."""
    t = np.asarray(times, np.float64)
    ok = np.isfinite(t)
    s = np.full(t.size, -np.inf)
    s[ok] = -tau * np.log(t[ok] / t[ok].min())
    e = np.where(ok, np.exp(s - s[ok].max()), 0.0)
    return (e / e.sum()).astype(np.float32)


def droptol_times(A: CSR, grid=DROPTOL_GRID, tol: float = 1e-8, seed: int = 77):
    """Seconds of factor set-up (fastainv_sym on the host) plus FP64 FSAI-PCG solve on the GPU
    for each log2 drop tolerance of the grid (solve timed after one warm-up solve)."""
    from .solver import PCG, Solver
    xt = G.uniform(seed, A.n_rows)
    b = np.asarray(A.to_scipy() @ xt)
    out = np.full(len(grid), np.inf)
    for k, lg in enumerate(grid):
        t0 = time.perf_counter()
        M = fastainv_sym(A, F64, max_elts=64 * A.n_rows, droptol=2.0 ** lg)
        setup = time.perf_counter() - t0
        sol = Solver(A, M=M, precision="fp64", method=PCG, tol=tol, maxit=20000, fsai_product=False)
        n = sol.n
        bd = DeviceArray.from_numpy(np.concatenate([b, np.zeros(n - A.n_rows)]))
        xd = DeviceArray(n)
        sol.solve_device(bd, xd, allow_fail=True)
        xd.zero()
        synchronize()
        t0 = time.perf_counter()
        st = sol.solve_device(bd, xd, allow_fail=True)
        synchronize()
        if st.converged:
            out[k] = setup + time.perf_counter() - t0
    return out


# ----------------------------------------------------------------------------- corpus --
def corpus_spec(i: int) -> dict:
    """Deterministic description of corpus matrix i: 7/19/27-point stencils (5e4..1e6 rows,
    z anisotropy 1..100 and coefficient contrast 1..1e6, all log-uniform) and random
    geometric graph Laplacians (3e4..3e5 rows, 10..26 neighbours)."""
    u = G.uniform01(900001 + i, 8)
    if u[0] < 0.75:
        pts = (7, 19, 27)[min(2, int(u[1] * 3))]
        n = int(10 ** (4.7 + u[2] * 1.3))
        nx = max(8, int(round(n ** (1 / 3) * (0.7 + 0.6 * u[3]))))
        ny = max(8, int(round(n ** (1 / 3) * (0.7 + 0.6 * u[4]))))
        nz = max(8, n // (nx * ny))
        return dict(kind="stencil", nx=nx, ny=ny, nz=nz, points=pts, anisotropy=float(10 ** (2 * u[5])),
                    contrast=float(10 ** (6 * u[6])), seed=1000 + i)
    return dict(kind="rgg", n=int(10 ** (4.5 + u[2])), k=10 + min(16, int(u[3] * 17)), seed=2000 + i)


def make_matrix(spec: dict) -> CSR:
    if spec["kind"] == "stencil":
        return G.stencil(spec["nx"], spec["ny"], spec["nz"], spec["points"], spec["anisotropy"],
                         spec["contrast"], seed=spec["seed"])
    return G.rgg_laplacian(spec["n"], spec["k"], spec["seed"])


def label(first: int, count: int, out: str, log=print, droptol=True):
    """Features and measured labels of corpus matrices first .. first+count-1, saved to `out`
    after every matrix (a partial file is resumed)."""
    X, T, DT, idx = [], [], [], []
    if os.path.exists(out):
        z = np.load(out, allow_pickle=False)
        X, T, DT, idx = list(z["X"]), list(z["T"]), list(z["DT"]), list(z["idx"])
    done = set(int(i) for i in idx)
    for i in range(first, first + count):
        if i in done:
            continue
        t0 = time.perf_counter()
        A = make_matrix(corpus_spec(i))
        X.append(features(A))
        T.append(time_formats(A))
        DT.append(droptol_times(A) if droptol else np.full(len(DROPTOL_GRID), np.nan))
        idx.append(i)
        np.savez(out, X=np.array(X, np.float32), T=np.array(T), DT=np.array(DT), idx=np.array(idx))
        dt = DT[-1]
        dbest = f"2^{DROPTOL_GRID[int(np.nanargmin(dt))]:.0f}" if np.isfinite(dt).any() else "n/a"
        log(f"matrix {i}: n={A.n_rows} nnz={A.nnz} best {CLASSES[int(np.argmin(T[-1]))]} "
            f"{np.min(T[-1]) * 1e6:.1f} us, droptol {dbest} [{time.perf_counter() - t0:.1f}s]")
    return out


def load_corpus(paths):
    X, T, DT, idx = [], [], [], []
    for p in paths:
        z = np.load(p, allow_pickle=False)
        X.append(z["X"])
        T.append(z["T"])
        DT.append(z["DT"])
        idx.append(z["idx"])
    idx = np.concatenate(idx)
    order = np.argsort(idx)
    return (np.concatenate(X)[order], np.concatenate(T)[order], np.concatenate(DT)[order], idx[order])


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

    def __init__(self, params=None, mean=None, std=None, seed=0, confidence=0.3):
        self.confidence = float(confidence)
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
        """(class probabilities B x C, log2 droptol B)."""
        z = self.forward(X)
        zc = z[:, :N_CLASSES].astype(np.float64)
        p = np.exp(zc - zc.max(1, keepdims=True))
        p /= p.sum(1, keepdims=True)
        return p, z[:, N_CLASSES]

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

    def save(self, path):
        np.savez(path, params=self.params, mean=self.mean, std=self.std, confidence=self.confidence,
                 classes=np.array(CLASSES, np.int32), droptol_grid=np.array(DROPTOL_GRID))

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        if not np.array_equal(z["classes"], np.array(CLASSES, np.int32)):
            raise ValueError("the model was trained on another class list")
        conf = float(z["confidence"]) if "confidence" in z.files else 0.3
        return cls(z["params"], z["mean"], z["std"], confidence=conf)


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
    """(droptol, source): the model's prediction inside the trained range, else 0.125."""
    model = model or load_default()
    if model is not None:
        _, lg = model.predict(features(A))
        lg = float(lg[0])
        if min(DROPTOL_GRID) - 0.5 <= lg <= max(DROPTOL_GRID) + 0.5:
            return 2.0 ** float(np.clip(lg, min(DROPTOL_GRID), max(DROPTOL_GRID))), "model"
    return DEFAULT_DROPTOL, "default"


def best_fixed_class(T) -> int:
    """The one class with the smallest geometric-mean time ratio to the best over a set: the
    non-learned baseline a single hand-picked format would give."""
    return int(np.argmin(np.exp(np.log(T / T.min(1, keepdims=True)).mean(0))))


def evaluate(model, X, T, DT, confidence=None, fixed_class=None):
    """Held-out quality: for the plan head, the time of the chosen class (model when
    confident, else exhaustive) over the exhaustive best, and of the model's top class alone
    (and of one fixed class, when given); for the droptol head, set-up + solve time at the grid
    point nearest the prediction over that at the default 0.125 and over the best grid point."""
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
    grid = np.array(DROPTOL_GRID)
    ok = np.isfinite(DT).all(1)
    k = np.abs(np.clip(lg, grid.min(), grid.max())[:, None] - grid[None, :]).argmin(1)
    t_pred = DT[np.arange(len(k)), k]
    t_def = DT[:, 0]                 # grid[0] = log2(0.125), the default
    if ok.any():
        t_best = DT[ok].min(1)
        res["droptol_pred_over_default_gmean"] = float(np.exp(np.mean(np.log(t_pred[ok] / t_def[ok]))))
        res["droptol_pred_over_best_gmean"] = float(np.exp(np.mean(np.log(t_pred[ok] / t_best))))
        res["droptol_default_over_best_gmean"] = float(np.exp(np.mean(np.log(t_def[ok] / t_best))))
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
    t = sub.add_parser("train")
    t.add_argument("--corpus", nargs="+", required=True)
    t.add_argument("--holdout", type=int, default=60)
    t.add_argument("--epochs", type=int, default=400)
    t.add_argument("--tau", type=float, default=8.0)
    t.add_argument("--lr", type=float, default=3e-3)
    t.add_argument("--out", default=MODEL_PATH)
    e = sub.add_parser("eval")
    e.add_argument("--corpus", nargs="+", required=True)
    e.add_argument("--holdout", type=int, default=60)
    e.add_argument("--model", default=MODEL_PATH)
    args = ap.parse_args(argv)
    from . import init
    init()

    def log(*x):
        print(*x, flush=True)
    if args.cmd == "label":
        label(args.first, args.count, args.out, log=log, droptol=not args.no_droptol)
        return 0
    X, T, DT, idx = load_corpus(args.corpus)
    ntr = X.shape[0] - args.holdout
    if args.cmd == "train":
        Tp, Mp, Td, Md = targets(T, DT, args.tau)
        model = Model(seed=0)
        model.fit(X[:ntr], Tp[:ntr], Mp[:ntr], Td[:ntr], Md[:ntr], epochs=args.epochs, lr=args.lr, log=log)
        model.calibrate(X[:ntr])
        model.save(args.out)
        fixed = best_fixed_class(T[:ntr])
        log("train", evaluate(model, X[:ntr], T[:ntr], DT[:ntr], fixed_class=fixed))
        log("holdout", evaluate(model, X[ntr:], T[ntr:], DT[ntr:], fixed_class=fixed))
        return 0
    model = Model.load(args.model)
    log("holdout", evaluate(model, X[ntr:], T[ntr:], DT[ntr:], fixed_class=best_fixed_class(T[:ntr])))
    return 0


if __name__ == "__main__":
    sys.exit(main())
