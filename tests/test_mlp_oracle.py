"""Pins the MLP oracle's backward pass against central finite differences (FP64, no rounding)."""
import numpy as np

import mlp as M


def _params(rng, scale=0.2):
    p = np.zeros(M.param_count())
    o = 0
    for i, n in zip(M.IN, M.OUT):
        w = rng.standard_normal((n, i)) * scale
        if n == 32:
            w[4:] = 0.0
        w[:, 33:] = 0.0 if i == 64 else w[:, 33:]
        p[o:o + n * i] = w.ravel()
        o += n * i
        p[o:o + n] = rng.standard_normal(n) * 0.1 * (1 if n != 32 else 1)
        if n == 32:
            p[o + 4:o + n] = 0.0
        o += n
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(rng)
    x = rng.standard_normal((8, 33))
    t = rng.standard_normal((8, 3))
    loss, g = M.loss_and_grads(p, x, t, bf16_on=False)
    out0, _ = M.forward(p, x, bf16_on=False)
    e2_fixed = ((out0[:, :3] - t) ** 2).sum(1)

    def loss_fixed(q):
        out, _ = M.forward(q, x, bf16_on=False)
        e2 = ((out[:, :3] - t) ** 2).sum(1)
        return float(np.mean(e2 / 3.0 + e2_fixed * np.exp(-out[:, 3]) / 3.0 + out[:, 3]))

    # probe a spread of parameters in every layer, including biases and the log-variance row
    idx = []
    o = 0
    for i, n in zip(M.IN, M.OUT):
        idx += [o + 3, o + 1 * i + 7, o + 3 * i + 32, o + n * i + 1]
        if n == 32:
            idx += [o + 3 * i + 5, o + n * i + 3]
        o += n * i + n
    for k in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (loss_fixed(pp) - loss_fixed(pm)) / (2 * h)
        assert abs(fd - g[k]) < 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159], np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # exactly halfway: ties go to the even mantissa
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-6
