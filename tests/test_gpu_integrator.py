"""GPU KDK integrator (src/integrator.rs:163-316) against the FP64 oracle and the reference's own
integrator tests (linear growth, step convergence, momentum)."""
import numpy as np
import pytest

from helpers import OMEGA_M, mode_amplitude, plane_wave

pytestmark = pytest.mark.gpu


def _pw(oracle, n, dtype=np.float64):
    return plane_wave(n, 1.0, 0.5, 0.02, lambda a: oracle.growth(a, dtype=np.float64),
                      lambda a: oracle.growth_rate(a, dtype=np.float64),
                      lambda a: oracle.cosmo_e(a, dtype=np.float64), dtype=dtype)


def _solver(treepm, n, box, split=None, cap=1 << 15):
    print(f"[gpu] KDK context mesh={n} box={box} split={split is not None} cap={cap}", flush=True)
    return treepm.Solver(n, box, OMEGA_M, split=split, max_particles=cap)


def test_linear_growth_matches_growth_factor(treepm, oracle):
    # src/integrator.rs:368-388: 32^3 particles, 32^3 mesh, a = 0.02 -> 0.1 in 40 steps, within 1%.
    pos, mom = _pw(oracle, 32)
    s = _solver(treepm, 32, 1.0, cap=32 ** 3)
    before = mode_amplitude(s.density_contrast(pos.astype(np.float32)), 32, 1.0)
    p1, m1 = s.run(pos, mom, 0.02, 0.1, 40)
    after = mode_amplitude(s.density_contrast(p1.astype(np.float32)), 32, 1.0)
    expected = oracle.growth(0.1, dtype=np.float64) / oracle.growth(0.02, dtype=np.float64)
    assert abs(after / before - expected) / expected < 0.01
    # and the trajectory matches the FP64 oracle's
    op, om = oracle.sim_run(pos, mom, 32, 1.0, 0.02, 0.1, 40, dtype=np.float64)
    d = np.abs(p1 - op)
    d = np.minimum(d, 1.0 - d)
    assert d.max() < 1e-5
    assert np.abs(m1 - om).max() < 1e-5 * np.abs(om).max()


def test_steps_converge_and_momentum_is_conserved(treepm, oracle):
    # src/integrator.rs:390-435
    pos, mom = _pw(oracle, 16)
    s = _solver(treepm, 32, 1.0, cap=16 ** 3)
    c, _ = s.run(pos, mom, 0.02, 0.1, 10)
    f, _ = s.run(pos, mom, 0.02, 0.1, 40)
    d = np.abs(c - f)
    d = np.minimum(d, 1.0 - d)
    assert d.max() < 1e-3
    _, m = s.run(pos, mom, 0.02, 0.1, 20)
    assert np.all(np.abs(m.sum(0)) < 1e-3 * np.abs(m[:, 0]).sum())


def test_treepm_kdk_matches_fp64_oracle(treepm, oracle):
    n = 16
    with treepm.Solver(n, 100.0, OMEGA_M, max_particles=n ** 3) as ic:
        pos, mom = ic.zeldovich(3, 2.0, 0.3 * 100.0 / n, mom_factor=0.02 ** 2 * 1.0, want_momenta=True)
    split = dict(treepm.force_split(100.0, n, n, 1 / 30, 0.0))
    s = _solver(treepm, n, 100.0, split, cap=n ** 3)
    p1, m1 = s.run(pos, mom, 0.02, 0.03, 3)
    op, om = oracle.sim_run(pos, mom, n, 100.0, 0.02, 0.03, 3, split=split, dtype=np.float64)
    d = np.abs(p1 - op)
    d = np.minimum(d, 100.0 - d)
    assert d.max() < 1e-6 * 100.0
    assert np.abs(m1 - om).max() < 1e-4 * np.abs(om).max()


def test_step_to_dev_matches_run(treepm, oracle):
    n = 16
    pos, mom = _pw(oracle, n)
    s = _solver(treepm, 16, 1.0, cap=n ** 3)
    ref_p, ref_m = s.run(pos, mom, 0.02, 0.04, 2)
    dp = treepm.DeviceArray.from_numpy(pos.ravel())
    dm = treepm.DeviceArray.from_numpy(mom.ravel())
    da = treepm.DeviceArray(3 * n ** 3, np.float32)
    s.accelerations_dev(dp.ptr, n ** 3, 0.02, da.ptr)
    a0 = 0.02
    for i in range(2):
        a1 = 0.02 * (2.0 ** ((i + 1) / 2))
        s.step_to_dev(dp.ptr, dm.ptr, n ** 3, a0, a1, da.ptr)
        a0 = a1
    s.sync()
    assert np.array_equal(dp.numpy().reshape(-1, 3), ref_p)
    assert np.array_equal(dm.numpy().reshape(-1, 3), ref_m)
