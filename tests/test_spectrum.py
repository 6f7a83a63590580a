"""The reference's IC spectrum (src/ics.rs:20-64, src/cosmology.rs:93-103) as computed by the
library's host code: BBKS transfer, sigma_8 normalisation, top-hat sigma(R)."""
import numpy as np
from scipy.integrate import quad


def test_bbks_transfer_is_one_at_large_scales(treepm):
    # src/ics.rs:186-190: T(1e-6) within 1e-3 of 1.
    k = 1e-6
    p, amp = treepm.power_spectrum(k)
    t2 = p / (amp * k ** treepm.N_S)
    assert abs(np.sqrt(t2) - 1.0) < 1e-3


def test_normalized_spectrum_gives_sigma8(treepm):
    # src/ics.rs:192-200: sigma(8) of the normalised spectrum is sigma8 to 1e-6.
    assert abs(treepm.power_sigma(8.0) - treepm.SIGMA8) / treepm.SIGMA8 < 1e-6


def _sigma_quad(treepm, radius):
    # sigma(R)^2 = (1 / 2 pi^2) int k^2 P(k) W(kR)^2 dk with adaptive quadrature over ln k.
    def integrand(lk):
        k = np.exp(lk)
        x = k * radius
        w = 1.0 if x < 1e-4 else 3.0 * (np.sin(x) - x * np.cos(x)) / x ** 3
        return k ** 3 * treepm.power_spectrum(k)[0] * w * w
    val = quad(integrand, np.log(1e-4), np.log(1e3), limit=400, epsabs=0, epsrel=1e-10)[0]
    return np.sqrt(val / (2 * np.pi ** 2))


def test_sigma_of_radius_matches_adaptive_quadrature(treepm):
    # The 8192-interval Simpson rule against adaptive quadrature of the same P(k) at other radii:
    # sigma falls with radius.
    prev = np.inf
    for r in (2.0, 4.0, 16.0, 32.0):
        s = treepm.power_sigma(r)
        assert abs(s - _sigma_quad(treepm, r)) < 1e-6 * s
        assert s < prev
        prev = s


def test_spectrum_shape(treepm):
    # P rises as k^n_s on large scales and turns over: the BBKS peak of k P(k)... P itself peaks
    # near k ~ 0.02 h/Mpc for Gamma = 0.21; at small scales P falls roughly as k^(n_s - 4) ln^2 k.
    ks = np.logspace(-4, 1, 200)
    p = np.array([treepm.power_spectrum(k)[0] for k in ks])
    kpk = ks[np.argmax(p)]
    assert 0.008 < kpk < 0.05
    slope_small = np.log(p[1] / p[0]) / np.log(ks[1] / ks[0])
    assert abs(slope_small - treepm.N_S) < 0.02
    assert treepm.power_spectrum(0.0)[0] == 0.0
