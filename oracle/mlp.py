"""CPU oracle of the learned short-range correction's network (test infrastructure).

Restates libtpm's MLP (csrc/mlp_kernels.hip) in numpy: 64 -> 128 -> 128 -> 128 -> 32 (4 used),
SiLU on hidden layers, y = first 3 outputs, s = 4th, loss
    L = (1/B) sum_b [ |e_b|^2 / 3 + sg(|e_b|^2) exp(-s_b) / 3 + s_b ],  e = y - t
(mean squared error for y; Gaussian negative log-likelihood for s with the error held constant,
so the gradient with respect to y is 2 e / (3 B) and with respect to s is (1 - |e|^2 exp(-s) / 3) / B). The parameter vector layout is the library's: per layer
Wt [out x in] then bias. `bf16=True` rounds to BF16 (round to nearest even) at the points where
the GPU path stores BF16 (inputs, weights, hidden activations, the output and hidden dZ); products
are accumulated in float64 here. `bf16=False` is the exact network, used to pin the analytic
backward pass against finite differences.
"""
import numpy as np

IN = (64, 128, 128, 128)
OUT = (128, 128, 128, 32)


def bf16(x):
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(params):
    params = np.asarray(params, np.float64)
    layers = []
    o = 0
    for i, n in zip(IN, OUT):
        wt = params[o:o + n * i].reshape(n, i)
        o += n * i
        b = params[o:o + n]
        o += n
        layers.append((wt, b))
    return layers


def param_count():
    return sum(n * i + n for i, n in zip(IN, OUT))


def _r(x, on):
    return bf16(x).astype(np.float64) if on else np.asarray(x, np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, x, bf16_on=True):
    """x: (B, 33) features -> (B, 4) outputs; returns (out, cache)."""
    x = np.asarray(x, np.float64)
    X = np.zeros((len(x), 64))
    X[:, :33] = x
    X = _r(X, bf16_on)
    cache = {"in": [X], "z": []}
    h = X
    layers = unpack(params)
    for li, (wt, b) in enumerate(layers):
        z = h @ _r(wt, bf16_on).T + b
        if bf16_on:
            z = z.astype(np.float32).astype(np.float64)
        cache["z"].append(z)
        if li < 3:
            h = _r(silu(z), bf16_on)
            cache["in"].append(h)
    return z[:, :4], cache


def loss_and_grads(params, x, t, bf16_on=True):
    """Loss and the gradient vector (library layout)."""
    t = np.asarray(t, np.float64)
    out, cache = forward(params, x, bf16_on)
    B = len(t)
    s = out[:, 3]
    sc = np.clip(s, -16.0, 16.0)  # exp() argument clamped as in the library
    es = np.exp(-sc)
    e = out[:, :3] - t
    e2 = (e ** 2).sum(1)
    loss = float(np.sum(e2 / 3.0 + e2 * es / 3.0 + s) / B)
    dz = np.zeros((B, 32))
    dz[:, :3] = 2.0 * e / 3.0 / B
    dz[:, 3] = (1.0 - np.where(s == sc, e2 * es / 3.0, 0.0)) / B
    dz = _r(dz, bf16_on)
    layers = unpack(params)
    grads = [None] * 4
    for li in range(3, -1, -1):
        wt, b = layers[li]
        xin = cache["in"][li]
        gw = dz.T @ xin
        gb = dz.sum(0)
        grads[li] = (gw, gb)
        if li == 0:
            break
        dy = dz @ _r(wt, bf16_on)
        if bf16_on:
            dy = dy.astype(np.float32).astype(np.float64)
        dz = _r(dy * silu_grad(cache["z"][li - 1]), bf16_on)
    flat = np.concatenate([np.concatenate([g.ravel(), gb]) for g, gb in grads])
    return loss, flat
