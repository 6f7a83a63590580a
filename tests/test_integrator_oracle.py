"""Pins the oracle's cosmology and KDK integrator (src/cosmology.rs, src/integrator.rs:163-316)
and the library's host-side KDK factors, against the reference's own tests, scipy quadrature and
closed forms. No GPU."""
import numpy as np
import pytest
from scipy.integrate import quad

from helpers import mode_amplitude, plane_wave


def test_cosmology_kats(oracle):
    # src/cosmology.rs:109-139
    assert abs(oracle.cosmo_e(1.0) - 1.0) < 1e-5
    assert abs(oracle.growth(1.0) - 1.0) < 1e-5
    c0 = dict(oracle.LCDM, omega_r=0.0)
    r1, r2 = oracle.growth(1e-3, c0) / 1e-3, oracle.growth(2e-3, c0) / 2e-3
    assert abs(r1 - r2) / r1 < 1e-4
    assert abs(oracle.growth(0.5) - 0.6068044) < 1e-4
    assert abs(oracle.growth(0.5, dtype=np.float64) - 0.6068044) < 1e-6


def test_growth_rate_is_dlnd_dlna(oracle):
    a, h = 0.3, 1e-4
    fd = (np.log(oracle.growth(a * np.exp(h), dtype=np.float64)) -
          np.log(oracle.growth(a * np.exp(-h), dtype=np.float64))) / (2 * h)
    assert abs(oracle.growth_rate(a, dtype=np.float64) - fd) < 1e-6


def _e(a, c):
    return np.sqrt(c["omega_r"] / a ** 4 + c["omega_m"] / a ** 3 + c["omega_l"])


@pytest.mark.parametrize("a0,a1", [(0.02, 0.025), (0.1, 0.2), (0.5, 0.51)])
def test_kick_drift_factors_match_quadrature(oracle, treepm, a0, a1):
    c = oracle.LCDM
    kick = quad(lambda a: 1 / (a * _e(a, c)), a0, a1, epsabs=0, epsrel=1e-13)[0]
    drift = quad(lambda a: 1 / (a ** 3 * _e(a, c)), a0, a1, epsabs=0, epsrel=1e-13)[0]
    ok = oracle.kick_factor(a0, a1, dtype=np.float64)
    od = oracle.drift_factor(a0, a1, dtype=np.float64)
    # 64-interval Simpson (src/integrator.rs:287-297) against adaptive quadrature
    assert abs(ok - kick) < 1e-7 * kick
    assert abs(od - drift) < 1e-7 * drift
    am, k1, d, k2 = treepm.kdk_factors(a0, a1)
    om = oracle.midpoint(a0, a1, dtype=np.float64)
    # the library's host factors are the same FP64 algorithm as the oracle's twin
    assert abs(am - om) < 1e-14
    assert abs(d - od) < 1e-13 * od
    assert abs(k1 - oracle.kick_factor(a0, om, dtype=np.float64)) < 1e-13 * kick
    assert abs(k2 - oracle.kick_factor(om, a1, dtype=np.float64)) < 1e-13 * kick
    # the midpoint splits the kick integral in half (src/integrator.rs:272-285)
    assert abs(k1 - k2) < 1e-7 * kick


def _plane_wave_inputs(oracle, n, amp, a_init, dtype):
    return plane_wave(n, 1.0, amp, a_init, lambda a: oracle.growth(a, dtype=dtype),
                      lambda a: oracle.growth_rate(a, dtype=dtype), lambda a: oracle.cosmo_e(a, dtype=dtype),
                      dtype=dtype)


def test_oracle_linear_growth_matches_growth_factor(oracle):
    # src/integrator.rs:368-388 on the FP32 oracle: 32^3 particles on a 32^3 mesh, a = 0.02 -> 0.1, 40 steps.
    pos, mom = _plane_wave_inputs(oracle, 32, 0.5, 0.02, np.float32)
    before = mode_amplitude(oracle.density_contrast(pos, 32, 1.0), 32, 1.0)
    pos, mom = oracle.sim_run(pos, mom, 32, 1.0, 0.02, 0.1, 40)
    after = mode_amplitude(oracle.density_contrast(pos, 32, 1.0), 32, 1.0)
    expected = oracle.growth(0.1) / oracle.growth(0.02)
    assert abs(after / before - expected) / expected < 0.01


def test_oracle_steps_converge_and_conserve_momentum(oracle):
    # src/integrator.rs:390-435
    p0, m0 = _plane_wave_inputs(oracle, 16, 0.5, 0.02, np.float32)
    c, _ = oracle.sim_run(p0, m0, 32, 1.0, 0.02, 0.1, 10)
    f, fm = oracle.sim_run(p0, m0, 32, 1.0, 0.02, 0.1, 40)
    d = np.abs(c.astype(np.float64) - f)
    d = np.minimum(d, 1.0 - d)
    assert d.max() < 1e-3
    p, m = oracle.sim_run(p0, m0, 32, 1.0, 0.02, 0.1, 20)
    scale = np.abs(m[:, 0]).sum()
    assert np.all(np.abs(m.astype(np.float64).sum(0)) < 1e-3 * scale)


@pytest.mark.parametrize("a", [0.02, 0.1, 0.5, 1.0])
def test_library_growth_matches_oracle(oracle, treepm, a):
    # tpm_growth is the host-side FP64 restatement of Cosmology::growth / growth_rate
    # (src/cosmology.rs:41-91); the oracle's FP64 twin is pinned by D(0.5) and f = dlnD/dlna above.
    d, f, e = treepm.growth(a)
    assert abs(d - oracle.growth(a, dtype=np.float64)) < 1e-12
    assert abs(f - oracle.growth_rate(a, dtype=np.float64)) < 1e-12
    assert abs(e - oracle.cosmo_e(a, dtype=np.float64)) < 1e-12 * e
    assert treepm.zeldovich_momentum_factor(a) == pytest.approx(a * a * f * e, rel=1e-15)


def test_growth_rate_matches_linder_today(treepm):
    # src/ics.rs:202-208: f(1) ~ Omega_m^0.55 to 0.01.
    _, f, _ = treepm.growth(1.0)
    assert abs(f - treepm.Cosmology.lcdm_default().omega_m ** 0.55) < 0.01
