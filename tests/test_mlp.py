"""Irradiance surrogate: the FP64 oracle pinned by finite differences and
closed forms (CPU), then the GPU's BF16-MFMA forward, backward and training
checked against it (GPU)."""
import numpy as np
import pytest

from oracle import mlp as M


def _points(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.0, 1.0, (n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return np.concatenate([p, nrm], 1).astype(np.float32)


def _smooth_target(x):
    """A smooth irradiance-like field of position and normal."""
    p, n = x[:, :3].astype(np.float64), x[:, 3:].astype(np.float64)
    base = 0.6 + 0.4 * np.sin(2.0 * p[:, 0] + 1.0) * np.cos(1.5 * p[:, 2]) + 0.3 * n[:, 1]
    return np.stack([base, 0.9 * base + 0.1 * p[:, 1], 1.1 * base - 0.05], 1).astype(np.float32)


# ---------------------------------------------------------------- CPU: oracle
def test_features_closed_form():
    x = np.zeros((1, 6), np.float32)
    f = M.features(x).reshape(6, 16, 2)
    assert np.all(f[..., 0] == 0.0) and np.all(f[..., 1] == 1.0)
    x = np.full((1, 6), 0.5, np.float32)
    f = M.features(x).reshape(6, 16, 2)
    # position: f_0 = 1: sin(pi/2) = 1; f_3 = 2: sin(pi) = 0, cos(pi) = -1
    assert np.allclose(f[:3, 0, 0], 1.0) and np.allclose(f[:3, 3, 1], -1.0)
    assert np.allclose(f[:3, 3, 0], 0, atol=1e-12)
    # normal: a quarter of the frequency: f_9 = 8 / 4 = 2
    assert np.allclose(f[3:, 9, 1], -1.0) and np.allclose(f[3:, 0, 0], np.sin(np.pi / 8))
    assert M.FREQ[15] == 32.0 and M.FREQ[9] == 8.0
    assert np.allclose(M.FREQ, 2.0 ** (np.arange(16) / 3.0), rtol=1e-7)


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = M.init_params(5).astype(np.float64)
    x = _points(9, 11)
    t = rng.uniform(0, 1, (9, 3))
    loss, g = M.loss_and_grad(params, x, t, emulate=False)
    idx = np.concatenate([rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 60, replace=False),
                          np.arange(M.OFF_B1, M.OFF_B1 + 8), rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 40),
                          rng.choice(np.arange(M.OFF_W3, M.OFF_B3), 40), np.arange(M.OFF_W4, M.N_PARAMS)])
    h = 1e-6
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, x, t, emulate=False)[0] - M.loss_and_grad(pm, x, t, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5], np.float32)
    got = M.bf16(vals)
    assert list(got) == [1.0, 1.0, 1.0 + 4 * 2 ** -8, 1.0, -2.5]


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_gpu_forward_matches_oracle(gpu):
    params = M.init_params(7)
    x = _points(1000, 1)
    net = gpu.Surrogate(params)
    y = net.infer(x)
    ref, _ = M.forward(params, x, emulate=True)
    err = np.abs(y - ref)
    scale = np.abs(ref).max()
    print(f"[mlp-fwd] max err {err.max():.3g} (scale {scale:.3g}), p99.9 {np.quantile(err, 0.999):.3g}", flush=True)
    # same BF16 operands; FP32 vs FP64 accumulation and the odd BF16 tie
    assert np.quantile(err, 0.999) <= 1e-4 * scale
    assert err.max() <= 1e-2 * scale
    assert np.array_equal(net.params(), params)


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle(gpu):
    params = M.init_params(9)
    x = _points(777, 2)  # not a multiple of the 32-point tile
    t = _smooth_target(x)
    net = gpu.Surrogate(params)
    loss, g = net.grad(x, t)
    rloss, rg = M.loss_and_grad(params, x, t, emulate=True)
    assert abs(loss - rloss) <= 1e-4 * rloss
    rel = np.linalg.norm(g - rg) / np.linalg.norm(rg)
    per = {}
    for name, a, b in (("W1", M.OFF_W1, M.OFF_B1), ("b1", M.OFF_B1, M.OFF_W2), ("W2", M.OFF_W2, M.OFF_B2),
                       ("W3", M.OFF_W3, M.OFF_B3), ("W4", M.OFF_W4, M.OFF_B4), ("b4", M.OFF_B4, M.N_PARAMS)):
        per[name] = float(np.linalg.norm(g[a:b] - rg[a:b]) / max(np.linalg.norm(rg[a:b]), 1e-30))
    print(f"[mlp-grad] loss gpu {loss:.6g} oracle {rloss:.6g}; rel grad err {rel:.3g} {per}", flush=True)
    assert rel <= 1e-3
    assert max(per.values()) <= 1e-2
    # BF16 operands barely turn the FP64 gradient
    _, g64 = M.loss_and_grad(params, x, t, emulate=False)
    cos = float(g @ g64 / (np.linalg.norm(g) * np.linalg.norm(g64)))
    assert cos > 0.99


@pytest.mark.gpu
def test_gpu_training_fits_a_smooth_field_deterministically(gpu):
    x = _points(8192, 4)
    t = _smooth_target(x)
    xv = _points(2048, 5)
    tv = _smooth_target(xv)
    base = float(np.mean((tv - t.mean(0)) ** 2) * 3)  # mean predictor
    a = gpu.Surrogate(seed=42)
    la = a.train(x, t, steps=300, batch=2048, lr=5e-3)
    b = gpu.Surrogate(seed=42)
    lb = b.train(x, t, steps=300, batch=2048, lr=5e-3)
    assert np.array_equal(a.params(), b.params()) and la == lb
    val = float(np.mean(np.sum((a.infer(xv) - tv) ** 2, 1)))
    print(f"[mlp-train] final batch loss {la:.4g}; held-out {val:.4g} vs mean predictor {base:.4g}", flush=True)
    assert val < 0.05 * base
