"""Learned short-range correction on the GPU: MFMA forward and backward against the CPU oracle
(oracle/mlp.py, BF16 rounding at the same points), training from walk targets, the error
estimator's fallback and what the trained network gains over the PM force alone."""
import numpy as np
import pytest

import mlp as M
from helpers import OMEGA_M, rel_errors

pytestmark = pytest.mark.gpu


def _random_params(rng):
    p = np.zeros(M.param_count(), np.float32)
    o = 0
    for i, n in zip(M.IN, M.OUT):
        w = (rng.standard_normal((n, i)) * np.sqrt(2.0 / i)).astype(np.float32)
        if n == 32:
            w[4:] = 0.0
        p[o:o + n * i] = w.ravel()
        o += n * i
        p[o:o + n] = (rng.standard_normal(n) * 0.1).astype(np.float32)
        if n == 32:
            p[o + 4:o + n] = 0.0
        o += n
    return p


def _solver(treepm, n=16):
    print(f"[gpu] mlp context mesh={n}", flush=True)
    split = treepm.force_split(100.0, n, n, 1 / 30, 0.35)
    return treepm.Solver(n, 100.0, OMEGA_M, split=split, max_particles=max(n ** 3, 70000))


def test_forward_matches_oracle(treepm):
    rng = np.random.default_rng(1)
    p = _random_params(rng)
    x = rng.standard_normal((100, 33)).astype(np.float32)
    s = _solver(treepm)
    s.mlp_set_params(p)
    out = s.mlp_predict(x)
    ref, _ = M.forward(p, x, bf16_on=True)
    assert np.max(np.abs(out - ref)) < 2e-2 * np.abs(ref).max()
    exact, _ = M.forward(p, x, bf16_on=False)
    assert np.max(np.abs(out - exact)) < 5e-2 * np.abs(exact).max()


def test_gradients_match_oracle(treepm):
    rng = np.random.default_rng(2)
    p = _random_params(rng)
    x = rng.standard_normal((256, 33)).astype(np.float32)
    t = rng.standard_normal((256, 3)).astype(np.float32)
    s = _solver(treepm)
    s.mlp_set_params(p)
    g, loss = s.mlp_gradients(x, t)
    rl, rg = M.loss_and_grads(p, x, t, bf16_on=True)
    assert abs(loss - rl) < 1e-3 * abs(rl)
    o = 0
    for i, n in zip(M.IN, M.OUT):
        for size in (n * i, n):
            a, b = g[o:o + size], rg[o:o + size]
            assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-7, (i, n, size)
            o += size


def test_training_reduces_loss_and_fallback_is_exact(treepm):
    n = 32
    s = _solver(treepm, n)
    pos, _ = s.zeldovich(5, 2.0, 1.5 * 100.0 / n)
    dp = treepm.DeviceArray.from_numpy(pos.ravel())
    acc = treepm.DeviceArray(3 * n ** 3, np.float32)
    s.mlp_init(3)
    hist = s.mlp_train_dev(dp.ptr, n ** 3, 0.2, n ** 3, epochs=60, batch=1024, lr=1e-3, seed=3)
    print(f"[gpu] training loss by epoch: {hist}", flush=True)
    assert hist[-1] < hist[0]
    # tol = 0 rejects every estimate: the force is the plain walk's, bit for bit.
    plain, _ = s.accelerations(pos, 0.2)
    nf = s.accelerations_learned_dev(dp.ptr, n ** 3, 0.2, 0.0, acc.ptr)
    s.sync()
    assert nf == n ** 3
    assert np.array_equal(acc.numpy().reshape(-1, 3).view(np.uint32), plain.view(np.uint32))
    # a huge tolerance accepts every estimate: PM + MLP, no walk. The trained network must beat
    # the PM force alone against the exact TreePM force.
    nf = s.accelerations_learned_dev(dp.ptr, n ** 3, 0.2, 1e30, acc.ptr)
    s.sync()
    assert nf == 0
    learned = acc.numpy().reshape(-1, 3)
    # The split PM force alone: the TreePM force minus its short-range part.
    k = 1.5 * OMEGA_M * 100.0 ** 3 / (4 * np.pi * 0.2 * n ** 3)
    pm_only = plain.astype(np.float64) - s.short_range(pos, k)
    e_learned = np.sqrt(np.mean(rel_errors(learned, plain) ** 2))
    e_pm = np.sqrt(np.mean(rel_errors(pm_only, plain) ** 2))
    print(f"[gpu] rms rel error vs TreePM: PM only {e_pm:.3e}, PM + MLP {e_learned:.3e}", flush=True)
    assert e_learned < e_pm
