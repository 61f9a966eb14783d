"""FP64 CPU ORACLE — TEST INFRASTRUCTURE ONLY.  Never imported by the
product path; only tests/, the smoke check and the benchmark's CPU leg use it.

Restatement of the irradiance surrogate (no reference counterpart: the
network stands in for RadiosityFunction::GatherLight,
core/lighting/radiosity.cpp:509-743):

    features  sin(pi f_jk x_j), cos(pi f_jk x_j), f_jk = 2^(k/3), k = 0..15,
              for the position (j = 0..2) and a quarter of that for the
              normal (j = 3..5)  (feature 32 j + 2 k + {0: sin, 1: cos})
    h1 = relu(F W1 + b1), h2 = relu(h1 W2 + b2), h3 = relu(h2 W3 + b3)
    y  = h3 W4 + b4
    L  = (1/n) sum_p |y_p - t_p|^2

`emulate=True` rounds to BF16 exactly where the GPU kernels hold BF16
operands (weights, features, hidden activations, dy, dz) and computes
everything else in FP64, so the GPU result must agree up to FP32
accumulation order.  `emulate=False` is the plain FP64 network, whose
gradient is pinned by central finite differences in the tests.
"""
import numpy as np

N_IN, N_HID, N_OUT, N_FREQ = 192, 64, 3, 16
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN * N_HID
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID * N_HID
OFF_W4 = OFF_B3 + N_HID
OFF_B4 = OFF_W4 + N_HID * N_OUT
N_PARAMS = OFF_B4 + N_OUT

# FP32 2^((k mod 3)/3), scaled exactly by 2^(k div 3)
FREQ = np.array([np.float32(2.0 ** ((k % 3) / 3.0)) * np.float32(2 ** (k // 3)) for k in range(N_FREQ)],
                np.float32)


def bf16(a):
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return dict(W1=p[OFF_W1:OFF_B1].reshape(N_IN, N_HID), b1=p[OFF_B1:OFF_W2],
                W2=p[OFF_W2:OFF_B2].reshape(N_HID, N_HID), b2=p[OFF_B2:OFF_W3],
                W3=p[OFF_W3:OFF_B3].reshape(N_HID, N_HID), b3=p[OFF_B3:OFF_W4],
                W4=p[OFF_W4:OFF_B4].reshape(N_HID, N_OUT), b4=p[OFF_B4:N_PARAMS])


def pack(d):
    return np.concatenate([d["W1"].ravel(), d["b1"], d["W2"].ravel(), d["b2"], d["W3"].ravel(),
                           d["b3"], d["W4"].ravel(), d["b4"]])


# per input dimension: position frequencies, then a quarter of them (exact) for the normal
FREQ_DIM = np.stack([FREQ] * 3 + [FREQ * np.float32(0.25)] * 3).astype(np.float32)  # (6, 16)


def features(x):
    """(n, 6) float32 inputs -> (n, 192) features (FP32 argument, FP64 sin/cos)."""
    x = np.asarray(x, np.float32).reshape(-1, 6)
    v = (x[:, :, None] * FREQ_DIM[None, :, :]).astype(np.float32).astype(np.float64)  # (n, 6, 16)
    f = np.stack([np.sin(np.pi * v), np.cos(np.pi * v)], axis=-1)                    # (n, 6, 16, 2)
    return f.reshape(x.shape[0], N_IN)


def forward(params, x, emulate=True):
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    P = unpack(params)
    W = {k: (r(v) if k.startswith("W") else v) for k, v in P.items()}
    F = r(features(x))
    h1 = r(np.maximum(F @ W["W1"] + W["b1"], 0.0))
    h2 = r(np.maximum(h1 @ W["W2"] + W["b2"], 0.0))
    h3 = r(np.maximum(h2 @ W["W3"] + W["b3"], 0.0))
    y = h3 @ W["W4"] + W["b4"]
    return y, dict(W=W, F=F, h1=h1, h2=h2, h3=h3)


def loss_and_grad(params, x, t, emulate=True):
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    t = np.asarray(t, np.float64).reshape(-1, 3)
    n = t.shape[0]
    y, c = forward(params, x, emulate)
    W = c["W"]
    e = y - t
    loss = float(np.sum(e * e) / n)
    dy = r(2.0 * e / n)
    dz3 = r((dy @ W["W4"].T) * (c["h3"] > 0))
    dz2 = r((dz3 @ W["W3"].T) * (c["h2"] > 0))
    dz1 = r((dz2 @ W["W2"].T) * (c["h1"] > 0))
    g = dict(W4=c["h3"].T @ dy, b4=dy.sum(0), W3=c["h2"].T @ dz3, b3=dz3.sum(0),
             W2=c["h1"].T @ dz2, b2=dz2.sum(0), W1=c["F"].T @ dz1, b1=dz1.sum(0))
    return loss, pack(g)


def init_params(seed=1):
    """Seeded He-uniform weights, zero biases (any fixed vector serves the
    parity tests; the library draws its own with splitmix64)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, fi, shape, gain in (("W1", N_IN, (N_IN, N_HID), 1.0), ("W2", N_HID, (N_HID, N_HID), 1.0),
                                  ("W3", N_HID, (N_HID, N_HID), 1.0), ("W4", N_HID, (N_HID, N_OUT), 0.5)):
        lim = gain * np.sqrt(6.0 / fi)
        d[name] = rng.uniform(-lim, lim, shape)
    d.update(b1=rng.uniform(-0.1, 0.1, N_HID), b2=rng.uniform(-0.1, 0.1, N_HID),
             b3=rng.uniform(-0.1, 0.1, N_HID), b4=rng.uniform(-0.1, 0.1, N_OUT))
    return pack(d).astype(np.float32)
