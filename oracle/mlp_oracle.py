"""mlp_oracle.py -- numpy restatement of the learned planner's network and loss (the HIP kernels
in ispm-sparse-lib_amd/csrc/mlp.hip).  TEST INFRASTRUCTURE ONLY.

Pinned in tests/test_mlp.py by central finite differences of the loss (which do not restate
the backward formulas) and by a hand-computed 1-sample case."""
import numpy as np

IN, H = 32, 64


def split(params, C):
    O = C + 1
    p = np.asarray(params, np.float64)
    o = 0
    out = {}
    for name, shape in (("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (O, H)), ("b3", (O,))):
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    assert o == p.size
    return out


def forward(params, X, C):
    P = split(params, C)
    h1 = np.maximum(X @ P["W1"].T + P["b1"], 0)
    h2 = np.maximum(h1 @ P["W2"].T + P["b2"], 0)
    z = h2 @ P["W3"].T + P["b3"]
    return z, (h1, h2)


def loss(params, X, Tp, Mp, Td, Md, C, wdrop):
    z, _ = forward(params, X, C)
    B = X.shape[0]
    zc = z[:, :C]
    lse = np.log(np.exp(zc - zc.max(1, keepdims=True)).sum(1)) + zc.max(1)
    lp = -(Tp * (zc - lse[:, None])).sum(1) * (Mp > 0)
    ld = wdrop * (z[:, C] - Td) ** 2 * (Md > 0)
    return float((lp + ld).sum() / B)


def gradient(params, X, Tp, Mp, Td, Md, C, wdrop):
    P = split(params, C)
    z, (h1, h2) = forward(params, X, C)
    B = X.shape[0]
    zc = z[:, :C]
    sm = np.exp(zc - zc.max(1, keepdims=True))
    sm /= sm.sum(1, keepdims=True)
    dz = np.zeros_like(z)
    dz[:, :C] = (sm - Tp) * (Mp > 0)[:, None] / B
    dz[:, C] = 2 * wdrop * (z[:, C] - Td) * (Md > 0) / B
    g = {}
    g["W3"] = dz.T @ h2
    g["b3"] = dz.sum(0)
    dh2 = (dz @ P["W3"]) * (h2 > 0)
    g["W2"] = dh2.T @ h1
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ P["W2"]) * (h1 > 0)
    g["W1"] = dh1.T @ X
    g["b1"] = dh1.sum(0)
    return np.concatenate([g[k].ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3")])
