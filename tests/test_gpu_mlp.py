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


def test_inference_kernel_gives_the_training_pass_loss(treepm):
    # The fused inference kernel (weights in LDS, activations exchanged across lanes) and the training
    # pass's forward (one GEMM per layer) do the same arithmetic: the loss of the inference outputs
    # equals the loss the training pass reports.
    rng = np.random.default_rng(4)
    p = _random_params(rng)
    x = rng.standard_normal((333, 33)).astype(np.float32)
    t = rng.standard_normal((333, 3)).astype(np.float32)
    s = _solver(treepm)
    s.mlp_set_params(p)
    out = s.mlp_predict(x).astype(np.float64)
    _, loss = s.mlp_gradients(x, t)
    e2 = ((out[:, :3] - t) ** 2).sum(1)
    sc = np.clip(out[:, 3], -16, 16)
    ours = float(np.mean(e2 / 3 + e2 * np.exp(-sc) / 3 + out[:, 3]))
    print(f"[gpu] loss from inference outputs {ours:.8e}, training pass {loss:.8e}", flush=True)
    assert abs(ours - loss) <= 1e-5 * abs(loss)


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


def _gate_check(treepm, kind):
    """Train on one realisation of a 64^3 workload and apply to another; returns the per-tolerance
    (tol, predicted fraction, errors of the predicted particles) after checking that every walked
    particle's force is the plain TreePM force bit for bit. With TPM_FLAG_REF_WALK the fallback is the
    per-particle walk, whose result for a particle does not depend on which others are walked."""
    n = 64
    npart = n ** 3
    a = 1.0 if kind == "nfw" else 0.02
    split = treepm.force_split(100.0, n, n, 1 / 30, 0.35)
    s = treepm.Solver(n, 100.0, OMEGA_M, split=split, max_particles=npart,
                      flags=treepm.FLAG_REF_WALK | treepm.FLAG_REF_LEAF_RULE)
    train = treepm.DeviceArray(3 * npart, np.float64)
    test = treepm.DeviceArray(3 * npart, np.float64)
    acc = treepm.DeviceArray(3 * npart, np.float32)
    walked = treepm.DeviceArray(npart, np.uint32)
    out = []
    try:
        for seed, dev in ((11, train), (12, test)):
            if kind == "nfw":
                s.nfw_halos_dev(treepm.NfwSpec.clustered(seed=seed, n_halos=400), npart, dev.ptr)
            else:
                s.zeldovich_dev(seed, 2.0, 0.3 * 100.0 / n, 0.0, dev.ptr)
        s.mlp_init(5)
        print(f"[gpu] {kind} gate: training", flush=True)
        hist = s.mlp_train_dev(train.ptr, npart, a, npart // 2, epochs=60, batch=1024, lr=1e-3, seed=5)
        print(f"[gpu] {kind}: loss {hist[0]:.3f} -> {hist[-1]:.3f}", flush=True)
        assert hist[-1] < hist[0]
        s.accelerations_dev(test.ptr, npart, a, acc.ptr)
        plain = acc.numpy().reshape(-1, 3).copy()
        for tol in (0.02, 0.05, 0.1, 0.2, 0.5):
            nf = s.accelerations_learned_dev(test.ptr, npart, a, tol, acc.ptr, walked.ptr)
            got = acc.numpy().reshape(-1, 3)
            mask = walked.numpy().astype(bool)
            assert int(mask.sum()) == nf
            assert np.array_equal(got[mask].view(np.uint32), plain[mask].view(np.uint32))
            assert np.isfinite(got).all()
            e = rel_errors(got[~mask], plain[~mask])
            frac = 1.0 - nf / npart
            print(f"[gpu] {kind} tol {tol}: kappa {s.mlp_error_scale(tol):.3g}, {frac:.4f} predicted, their error "
                  f"p50 {np.median(e) if len(e) else 0:.3e} p95 {np.percentile(e, 95) if len(e) else 0:.3e}",
                  flush=True)
            out.append((tol, frac, e))
    finally:
        s.close()
        for d in (train, test, acc, walked):
            d.free()
    return out


def test_learned_gate_meets_tolerance(treepm):
    # The lattice Zel'dovich workload (z = 49, 0.3-cell displacements): the gate is calibrated so that
    # 99% of the held-out particles it accepts are within tol, so on another realisation 95% must be.
    # Checked at the tightest listed tolerance at which the network takes at least 1% of the particles.
    res = _gate_check(treepm, "zeldovich")
    chosen = [(tol, e) for tol, frac, e in res if frac >= 0.01]
    assert chosen, "the network takes no particle at any listed tolerance"
    tol, e = chosen[0]
    assert np.percentile(e, 95) <= tol


def test_clustered_fallback_is_exact(treepm):
    # The clustered workload at reduced size (NFW halos + 30% background): the network takes few or no
    # particles at these tolerances, every walked one is exact, and the few it takes stay within the
    # loosest tolerance at the 95th percentile.
    res = _gate_check(treepm, "nfw")
    tol, frac, e = res[-1]
    if len(e) >= 20:
        assert np.percentile(e, 95) <= tol
