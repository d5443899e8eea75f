"""The learned planner's network: the numpy oracle pinned by finite differences and a hand
case (CPU), the HIP forward / gradient / Adam kernels against the oracle (GPU), and the
planner's decisions and fallbacks."""
import os

import numpy as np
import pytest

import ispm_amd as I
import ispm_amd.gen as G
from ispm_amd import planner as PL
import mlp_oracle as MO

C_ = PL.N_CLASSES


def random_case(B=6, seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    n = MO.IN * MO.H + MO.H + MO.H * MO.H + MO.H + (C_ + 1) * MO.H + C_ + 1
    params = rng.normal(0, scale, n)
    X = rng.normal(0, 1, (B, MO.IN))
    Tp = rng.random((B, C_))
    Tp /= Tp.sum(1, keepdims=True)
    Mp = (rng.random(B) < 0.8).astype(float)
    Td = rng.normal(-5, 1, B)
    Md = (rng.random(B) < 0.7).astype(float)
    return params, X, Tp, Mp, Td, Md


def test_oracle_gradient_matches_finite_differences():
    params, X, Tp, Mp, Td, Md = random_case()
    g = MO.gradient(params, X, Tp, Mp, Td, Md, C_, 0.5)
    rng = np.random.default_rng(1)
    idx = rng.choice(params.size, 60, replace=False)
    h = 1e-6
    for k in idx:
        pp, pm = params.copy(), params.copy()
        pp[k] += h
        pm[k] -= h
        fd = (MO.loss(pp, X, Tp, Mp, Td, Md, C_, 0.5) - MO.loss(pm, X, Tp, Mp, Td, Md, C_, 0.5)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_oracle_hand_case():
    """All weights zero: z = b3, so the loss is the cross-entropy of softmax(b3[:C]) plus
    w (b3[C] - Td)^2, and d loss / d b3 = softmax - q and 2 w (b3[C] - Td)."""
    n = MO.IN * MO.H + MO.H + MO.H * MO.H + MO.H + (C_ + 1) * MO.H + C_ + 1
    params = np.zeros(n)
    b3 = np.linspace(-1, 1, C_ + 1)
    params[-(C_ + 1):] = b3
    X = np.ones((1, MO.IN))
    q = np.zeros((1, C_))
    q[0, 3] = 1
    Td = np.array([-4.0])
    one = np.ones(1)
    sm = np.exp(b3[:C_]) / np.exp(b3[:C_]).sum()
    want = -np.log(sm[3]) + 0.5 * (b3[C_] + 4.0) ** 2
    assert MO.loss(params, X, q, one, Td, one, C_, 0.5) == pytest.approx(want, rel=1e-12)
    g = MO.gradient(params, X, q, one, Td, one, C_, 0.5)
    gb3 = g[-(C_ + 1):]
    np.testing.assert_allclose(gb3[:C_], sm - q[0], rtol=1e-12, atol=1e-15)
    assert gb3[C_] == pytest.approx(2 * 0.5 * (b3[C_] + 4.0))
    assert np.all(g[:-(C_ + 1)] == 0)      # zero weights: ReLU outputs 0, no other gradient


def test_plan_target_and_targets():
    t = np.array([2.0, 1.0, np.inf, 1.5] + [3.0] * (C_ - 4))
    q = PL.plan_target(t)
    assert q.sum() == pytest.approx(1.0, rel=1e-6)
    assert q.argmax() == 1 and q[2] == 0
    DT = np.array([[1.0, 0.5, 0.7, np.inf, 0.9]])
    Tp, Mp, Td, Md = PL.targets(t[None, :], DT)
    assert Mp[0] == 1 and Md[0] == 1 and Td[0] == PL.DROPTOL_GRID[1]
    Tp, Mp, Td, Md = PL.targets(t[None, :], np.full((1, 5), np.inf))
    assert Md[0] == 0


def test_features_known_values():
    A = G.poisson7(8)
    f = PL.features(A)
    assert f.shape == (32,) and np.isfinite(f).all()
    rl = np.diff(A.row_ptr)
    assert f[2] == pytest.approx(rl.mean() / 32)
    assert f[18] == pytest.approx(1.0)              # every entry on one of 7 diagonals
    assert f[31] == 1.0


def test_corpus_specs_deterministic():
    a = [PL.corpus_spec(i) for i in range(20)]
    b = [PL.corpus_spec(i) for i in range(20)]
    assert a == b
    kinds = {s["kind"] for s in a}
    assert kinds == {"stencil", "rgg"}


# ------------------------------------------------------------------------------ GPU --
@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle(gpu):
    params, X, Tp, Mp, Td, Md = random_case(B=37, seed=3)
    p32 = params.astype(np.float32)
    dev = {k: I.DeviceArray.from_numpy(np.asarray(v, np.float32).ravel())
           for k, v in dict(p=p32, X=X, Tp=Tp, Mp=Mp, Td=Td, Md=Md).items()}
    B = X.shape[0]
    out = I.DeviceArray(B * (C_ + 1), np.float32)
    I.check(I.lib().ispm_mlp_forward(C_, dev["p"].vp, dev["X"].vp, B, out.vp))
    z_ref, _ = MO.forward(p32.astype(np.float64), X.astype(np.float32).astype(np.float64), C_)
    z = out.numpy().reshape(B, C_ + 1)
    np.testing.assert_allclose(z, z_ref, rtol=1e-4, atol=1e-4 * np.abs(z_ref).max())
    grad = I.DeviceArray(params.size, np.float32)
    loss = I.DeviceArray(1, np.float32)
    I.check(I.lib().ispm_mlp_gradient(C_, dev["p"].vp, dev["X"].vp, dev["Tp"].vp, dev["Mp"].vp, dev["Td"].vp,
                                      dev["Md"].vp, B, I.C.c_float(0.5), grad.vp, loss.vp))
    args = [np.asarray(a, np.float32).astype(np.float64) for a in (X, Tp, Mp, Td, Md)]
    g_ref = MO.gradient(p32.astype(np.float64), *args, C_, 0.5)
    g = grad.numpy()
    np.testing.assert_allclose(g, g_ref, rtol=2e-3, atol=2e-4 * np.abs(g_ref).max())
    assert loss.numpy()[0] == pytest.approx(MO.loss(p32.astype(np.float64), *args, C_, 0.5), rel=1e-4)


@pytest.mark.gpu
def test_gpu_adam_steps_follow_numpy_and_reduce_loss(gpu):
    params, X, Tp, Mp, Td, Md = random_case(B=32, seed=4, scale=0.2)
    p32 = params.astype(np.float32)
    args32 = [np.asarray(a, np.float32) for a in (X, Tp, Mp, Td, Md)]
    dargs = [I.DeviceArray.from_numpy(a.ravel()) for a in args32]
    pd = I.DeviceArray.from_numpy(p32)
    m = I.DeviceArray(p32.size, np.float32)
    v = I.DeviceArray(p32.size, np.float32)
    loss = I.DeviceArray(1, np.float32)
    # numpy Adam on the oracle gradient for 3 steps
    pn, mn, vn = p32.astype(np.float64), np.zeros(p32.size), np.zeros(p32.size)
    a64 = [a.astype(np.float64) for a in args32]
    lr = 1e-3
    for step in range(1, 4):
        I.check(I.lib().ispm_mlp_train_step(C_, pd.vp, m.vp, v.vp, step, *[d.vp for d in dargs], 32,
                                            I.C.c_float(lr), I.C.c_float(0.5), loss.vp))
        g = MO.gradient(pn, *a64, C_, 0.5)
        mn = 0.9 * mn + 0.1 * g
        vn = 0.999 * vn + 0.001 * g * g
        pn = pn - lr * (mn / (1 - 0.9 ** step)) / (np.sqrt(vn / (1 - 0.999 ** step)) + 1e-8)
    np.testing.assert_allclose(pd.numpy(), pn, rtol=0, atol=5e-5)
    l0 = MO.loss(p32.astype(np.float64), *a64, C_, 0.5)
    for step in range(4, 300):
        I.check(I.lib().ispm_mlp_train_step(C_, pd.vp, m.vp, v.vp, step, *[d.vp for d in dargs], 32,
                                            I.C.c_float(3e-3), I.C.c_float(0.5), loss.vp))
    l1 = MO.loss(pd.numpy().astype(np.float64), *a64, C_, 0.5)
    assert l1 < 0.5 * l0


@pytest.mark.gpu
def test_gpu_model_fit_predict_and_fallbacks(gpu, tmp_path):
    rng = np.random.default_rng(5)
    N = 96
    X = rng.normal(0, 1, (N, 32)).astype(np.float32)
    # a learnable rule: class by the sign pattern of two features, droptol linear in a third
    cls = (X[:, 0] > 0).astype(int) * 2 + (X[:, 1] > 0).astype(int)
    T = np.full((N, C_), 2.0)
    T[np.arange(N), cls] = 1.0
    DT = np.full((N, len(PL.DROPTOL_GRID)), 2.0)
    dk = np.clip((X[:, 2] + 2) * 1.2, 0, len(PL.DROPTOL_GRID) - 1).astype(int)
    DT[np.arange(N), dk] = 1.0
    Tp, Mp, Td, Md = PL.targets(T, DT)
    model = PL.Model(seed=1)
    hist = model.fit(X, Tp, Mp, Td, Md, epochs=150, batch=32, lr=3e-3)
    assert hist[-1] < hist[0]
    p, lg = model.predict(X)
    assert (p.argmax(1) == cls).mean() > 0.9
    path = tmp_path / "m.npz"
    model.save(path)
    m2 = PL.Model.load(path)
    np.testing.assert_array_equal(m2.params, model.params)
    # decisions: confident model, or the exhaustive table when the threshold is unreachable
    A = G.poisson7(16)
    s = PL.choose_plan(A, model=m2, confidence=1.01)
    assert s[-1] == "exhaustive" and s[:4] in PL.CLASSES
    d, src = PL.choose_droptol(A, model=m2)
    assert src in ("model", "default") and 2.0 ** min(PL.DROPTOL_GRID) <= d <= 0.125


@pytest.mark.gpu
def test_gpu_time_formats_table(gpu):
    A = G.stencil(16, 16, 16, 27, seed=9)
    t = PL.time_formats(A)
    assert t.shape == (C_,)
    ok = np.isfinite(t)
    assert ok.sum() >= C_ - 3 and (t[ok] > 0).all()


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(PL.MODEL_PATH), reason="no trained model in the package")
def test_gpu_committed_model_predicts(gpu):
    model = PL.load_default()
    A = G.stencil(20, 20, 20, 19, anisotropy=10, contrast=100, seed=11)
    p, lg = model.predict(PL.features(A))
    assert p.shape == (1, C_) and np.isfinite(p).all() and np.isfinite(lg).all()


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(PL.MODEL_PATH), reason="no trained model in the package")
def test_gpu_committed_model_beats_fixed_format_on_holdout(gpu):
    """What the trained network gains, on the held-out end of the committed corpus (matrices it
    was not trained on): the SpMV time of its top format over the exhaustive best, against the
    single format that is best on average over the training part (the method without the
    network picks one format for all matrices; the exhaustive search pays for 39 conversions
    and timings per matrix)."""
    data = os.path.join(os.path.dirname(PL.MODEL_PATH), "planner_data")
    paths = sorted(os.path.join(data, f) for f in os.listdir(data) if f.endswith(".npz"))
    X, T, DT, idx = PL.load_corpus(paths)
    hold = 50
    model = PL.load_default()
    res = PL.evaluate(model, X[-hold:], T[-hold:], DT[-hold:],
                      fixed_class=PL.best_fixed_class(T[:-hold]))
    assert res["plan_top_ratio_gmean"] < res["plan_fixed_ratio_gmean"]
    assert res["plan_top_ratio_gmean"] <= 1.05
    assert res["plan_choice_ratio_gmean"] <= res["plan_top_ratio_gmean"] + 1e-12
