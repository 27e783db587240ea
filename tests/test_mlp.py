"""The beat surrogate: the FP64 numpy reference pinned by finite differences
and Adam's closed form (CPU), then the gfx950 BF16-MFMA kernels against it
(GPU)."""
import numpy as np
import pytest

from tests import mlp_reference as ref
from ryr_hyperactivity_amd.surrogate import NPARAM, init_params, unpack

LIVE = np.zeros(NPARAM, dtype=bool)
_v = unpack(LIVE)
_v["W1"][:11, :] = True
_v["b1"][:] = True
_v["W2"][:] = True
_v["b2"][:] = True
_v["W3"][:, :2] = True
_v["b3"][:2] = True


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 11))
    y = (x[:, 0] + 0.5 * x[:, 1] - x[:, 3] > 0.3).astype(np.float64)
    t = np.sin(x[:, 0]) + 0.5 * x[:, 2]
    mask = (x[:, 4] > -1.0).astype(np.float64)
    return x, y, t, mask


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = init_params(5).astype(np.float64) * 1.5
    p[LIVE] += rng.normal(scale=0.05, size=LIVE.sum())          # non-zero biases too
    x, y, t, mask = _data(40, 1)
    s = rng.uniform(0.5, 2.0, size=40)
    _, g = ref.loss_and_grad(p, x, y, t, mask, s, w_apd=0.7)
    assert np.all(g[~LIVE] == 0.0)
    idx = rng.choice(np.flatnonzero(LIVE), size=60, replace=False)
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (ref.loss_and_grad(pp, x, y, t, mask, s, 0.7)[0] - ref.loss_and_grad(pm, x, y, t, mask, s, 0.7)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_reference_gelu_derivative_and_adam_first_step():
    xs = np.linspace(-6, 6, 241)
    g, dg = ref.gelu(xs)
    h = 1e-6
    fd = (ref.gelu(xs + h)[0] - ref.gelu(xs - h)[0]) / (2 * h)
    assert np.allclose(dg, fd, atol=1e-8)
    assert np.allclose(g[xs > 5], xs[xs > 5], rtol=1e-6) and np.all(np.abs(g[xs < -5]) < 1e-5)
    # first Adam step: m_hat = g, v_hat = g^2, so the update is lr * g / (|g| + eps)
    gr = np.array([0.3, -2.0, 1e-3, 0.0])
    p1, _, _ = ref.adam_step(np.zeros(4), gr, np.zeros(4), np.zeros(4), 1, lr=0.01)
    assert np.allclose(p1, -0.01 * gr / (np.abs(gr) + 1e-8))


@pytest.mark.gpu
def test_forward_matches_fp64_reference_within_bf16_budget():
    from ryr_hyperactivity_amd.surrogate import BeatMLP
    print("case mlp forward", flush=True)
    p = init_params(7) * 1.3
    p[LIVE & (np.arange(NPARAM) >= 2048)] += 0.02
    x, _, _, _ = _data(1000, 2)
    with BeatMLP(params=p) as m:
        prob, apd = m.forward(x)
    z3, _ = ref.forward(p.astype(np.float64), x)
    pr = 1 / (1 + np.exp(-z3[:, 0]))
    # BF16 operands (8-bit mantissa) in three chained products with FP32
    # accumulation: a few 2^-8 relative roundings per layer
    assert np.max(np.abs(prob - pr)) < 0.03 and np.mean(np.abs(prob - pr)) < 0.006
    assert np.max(np.abs(apd - z3[:, 1])) < 0.05 * (1 + np.max(np.abs(z3[:, 1])))
    assert np.mean(np.abs(apd - z3[:, 1])) < 0.01 * (1 + np.mean(np.abs(z3[:, 1])))


@pytest.mark.gpu
def test_gradient_matches_fp64_reference_within_bf16_budget():
    from ryr_hyperactivity_amd.surrogate import BeatMLP
    print("case mlp gradient", flush=True)
    rng = np.random.default_rng(9)
    p = init_params(11)
    x, y, t, mask = _data(333, 4)                   # not a multiple of 16: a ragged last tile
    s = rng.uniform(0.2, 1.0, size=333)
    with BeatMLP(params=p) as m:
        loss, g = m.grad(x, y, t, mask, s, w_apd=0.5)
    lr_, gr = ref.loss_and_grad(p.astype(np.float64), x, y, t, mask, s, 0.5)
    assert abs(loss - lr_) < 0.01 * abs(lr_)
    assert np.all(g[~LIVE] == 0.0)
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    cos = g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr))
    assert rel < 0.03 and cos > 0.999, (rel, cos)
    for blk, (a, b) in {"W1": (0, 2048), "W2": (2112, 6208), "W3": (6272, 7296)}.items():
        r = np.linalg.norm(g[a:b] - gr[a:b]) / np.linalg.norm(gr[a:b])
        assert r < 0.05, (blk, r)


@pytest.mark.gpu
def test_adam_step_matches_reference_on_the_kernel_gradient():
    from ryr_hyperactivity_amd.surrogate import BeatMLP
    print("case mlp adam", flush=True)
    p = init_params(13)
    x, y, t, mask = _data(256, 6)
    with BeatMLP(params=p) as m:
        _, g = m.grad(x, y, t, mask)
        m.train(x, y, t, mask, steps=1, batch=256, lr=1e-3, seed=1)
        p1 = m.params()
    exp, _, _ = ref.adam_step(p.astype(np.float64), g.astype(np.float64), 0, 0, 1, lr=1e-3)
    big = np.abs(g) > 1e-4 * np.abs(g).max()
    # the training pass adds the same samples in a permuted order: only the
    # rounding of the gradient differs, and the first step is lr*sign(g)
    assert np.max(np.abs(p1[big] - exp[big])) < 2e-5
    assert np.all(p1[~LIVE] == 0.0)


@pytest.mark.gpu
def test_training_on_gpu_learns_a_known_rule():
    from ryr_hyperactivity_amd.surrogate import BeatMLP
    print("case mlp training", flush=True)
    x, y, t, mask = _data(4096, 8)
    xv, yv, tv, mv = _data(1024, 9)
    with BeatMLP(seed=3) as m:
        l0, _ = m.grad(x, y, t, mask)
        trace = m.train(x, y, t, mask, steps=1500, batch=128, lr=3e-3, seed=2)
        prob, apd = m.forward(xv)
        p = m.params()
    assert np.all(np.isfinite(trace)) and trace[-100:].mean() < 0.2 * l0
    acc = np.mean((prob > 0.5) == (yv > 0.5))
    rmse = np.sqrt(np.sum(mv * (apd - tv) ** 2) / mv.sum())
    assert acc > 0.93 and rmse < 0.2, (acc, rmse)
    assert np.all(p[~LIVE] == 0.0)
