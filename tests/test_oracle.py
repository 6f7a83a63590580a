"""Pins the CPU oracle (a restatement of the reference TreePM path) before it is trusted.

Each check compares the oracle with something that fixes the answer independently of it: the
reference's own known-answer tests (src/mesh.rs:101-183, src/poisson.rs:114-165,
src/tree.rs:382-430, src/integrator.rs:437-461), golden vectors written by numpy/scipy in
FP64 (tests/golden/), closed forms and invariants.
"""
import os

import numpy as np
import pytest

from helpers import OMEGA_M, fixed_point_delta32, lattice, max_diff, max_norm, mesh_scattered, tree_scattered

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- CIC (src/mesh.rs:101-183)

def test_deposit_conserves_total_mass(oracle):
    g = oracle.deposit(mesh_scattered(5000, 1.0), 16, 1.0, 0.25)
    assert abs(g.sum(dtype=np.float64) - 0.25 * 5000) < 1e-5 * 0.25 * 5000


def test_deposit_matches_fp64_golden(oracle):
    z = np.load(os.path.join(GOLD, "cic_5000.npz"))
    g32 = oracle.deposit(z["pos"], 16, 1.0, 0.25)
    g64 = oracle.deposit(z["pos"], 16, 1.0, 0.25, dtype=np.float64)
    ref = z["grid"]
    scale = np.maximum(np.abs(ref), ref.mean())
    assert np.max(np.abs(g64 - ref) / scale) < 1e-12
    assert np.max(np.abs(g32 - ref) / scale) < 1e-5


def test_fixed_point_delta_helper_matches_golden():
    # The test helper that predicts the GPU's fixed-point density contrast bit for bit: against the
    # numpy FP64 golden grid (FP32 weights, so to FP32 accuracy), exact mass conservation, and a
    # uniform lattice giving delta = 0 exactly.
    z = np.load(os.path.join(GOLD, "cic_5000.npz"))
    d = fixed_point_delta32(z["pos"], 16, 1.0)
    ref = z["grid"] / (0.25 * 5000 / 16 ** 3) - 1.0
    assert np.max(np.abs(d - ref) / np.maximum(np.abs(ref + 1), 1.0)) < 1e-6
    assert abs(((d.astype(np.float64) + 1).sum() - 16 ** 3)) < 1e-3
    assert np.all(fixed_point_delta32(lattice(8, 1.0), 8, 1.0) == 0.0)


def test_particle_at_cell_centre_fills_one_cell(oracle):
    n, h = 8, 1.0 / 8
    g = oracle.deposit([[2.5 * h, 3.5 * h, 4.5 * h]], n, 1.0).reshape(n, n, n)
    assert abs(g[2, 3, 4] - 1.0) < 1e-5
    assert (g > 1e-12).sum() == 1


def test_lattice_deposits_uniform_grid(oracle):
    g = oracle.deposit(lattice(8, 2.0), 8, 2.0, 3.0)
    assert np.all(np.abs(g - 3.0) < 1e-5)


def test_particle_on_boundary_wraps(oracle):
    n, h = 4, 0.25
    g = oracle.deposit([[1.0, 0.5 * h, 0.5 * h]], n, 1.0).reshape(n, n, n)
    assert abs(g[n - 1, 0, 0] - 0.5) < 1e-5 and abs(g[0, 0, 0] - 0.5) < 1e-5


def test_interpolation_constant_and_centre(oracle):
    v = oracle.interpolate(np.full(512, 2.5, np.float32), 8, 1.0, mesh_scattered(200, 1.0))
    assert np.all(np.abs(v - 2.5) < 1e-5)
    cells = np.zeros((8, 8, 8), np.float32)
    cells[5, 1, 6] = 7.0
    h = 1.0 / 8
    assert abs(oracle.interpolate(cells.ravel(), 8, 1.0, [[5.5 * h, 1.5 * h, 6.5 * h]])[0] - 7.0) < 1e-5


# ----------------------------------------------------------------------- FFT (src/fft.rs)

@pytest.mark.parametrize("n", [6, 8, 16, 47, 94])
def test_fft_line_matches_numpy_golden(oracle, n):
    z = np.load(os.path.join(GOLD, "fft.npz"))
    x, ref = z[f"x{n}"], z[f"f{n}"]
    assert np.max(np.abs(oracle.fft1d(x, -1, np.float64) - ref)) < 1e-12 * np.abs(ref).max()
    assert np.max(np.abs(oracle.fft1d(x, -1, np.float32) - ref)) < 3e-6 * np.abs(ref).max()
    # inverse is unnormalised (src/fft.rs:32-35)
    assert np.max(np.abs(oracle.fft1d(ref, +1, np.float64) - n * x)) < 1e-11 * n * np.abs(x).max()


def test_fft_line_matches_direct_dft(oracle):
    # A mixed radix the golden set does not hold: 2*3*5*7.
    n = 210
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    t = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(t, t) / n) @ x
    assert np.max(np.abs(oracle.fft1d(x, -1, np.float64) - dft)) < 1e-10 * np.abs(dft).max()


def test_fft3_matches_numpy_golden(oracle):
    z = np.load(os.path.join(GOLD, "fft.npz"))
    out = oracle.fft3(z["g12"], -1, np.float64).reshape(12, 12, 12)
    assert np.max(np.abs(out - z["G12"])) < 1e-11 * np.abs(z["G12"]).max()


def test_wavenumber_nyquist_is_zero(oracle):
    L = 100.0
    assert oracle.wavenumber(8, 16, L) == 0.0
    assert abs(oracle.wavenumber(3, 16, L) - 3 * 2 * np.pi / L) < 1e-5
    assert abs(oracle.wavenumber(13, 16, L) + 3 * 2 * np.pi / L) < 1e-5
    assert oracle.wavenumber(2, 5, L) > 0 and oracle.wavenumber(3, 5, L) < 0


# ------------------------------------------------------------- Poisson (src/poisson.rs)

def _mesh_from_fn(n, f, dtype=np.float32):
    x = (np.arange(n) + 0.5) / n
    return np.broadcast_to(f(x)[:, None, None], (n, n, n)).astype(dtype).ravel()


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-5), (np.float64, 1e-12)])
def test_plane_wave_force_matches_analytic(oracle, dtype, tol):
    n, a = 16, 0.5
    k0 = 2 * np.pi
    c = 1.5 * OMEGA_M / a
    g = oracle.poisson_accel(_mesh_from_fn(n, lambda x: np.cos(k0 * x), dtype), n, 1.0, OMEGA_M, a, 0.0, dtype=dtype)
    exp = _mesh_from_fn(n, lambda x: -(c / k0) * np.sin(k0 * x), np.float64)
    amp = c / k0
    assert np.max(np.abs(g[0] - exp)) < tol * amp
    assert np.max(np.abs(g[1])) < tol * amp and np.max(np.abs(g[2])) < tol * amp


def test_plane_wave_with_split_filter(oracle):
    # The filter multiplies the single mode by exp(-k0^2 r_s^2).
    n, a, rs = 16, 1.0, 0.05
    k0 = 2 * np.pi
    g = oracle.poisson_accel(_mesh_from_fn(n, lambda x: np.cos(k0 * x), np.float64), n, 1.0, OMEGA_M, a, rs,
                             dtype=np.float64)
    amp = 1.5 * OMEGA_M / a / k0 * np.exp(-k0 * k0 * rs * rs)
    exp = _mesh_from_fn(n, lambda x: -amp * np.sin(k0 * x), np.float64)
    assert np.max(np.abs(g[0] - exp)) < 1e-12 * amp


def test_uniform_delta_gives_no_force(oracle):
    g = oracle.poisson_accel(np.ones(512, np.float32), 8, 1.0, OMEGA_M, 1.0, 0.0)
    assert np.max(np.abs(g)) < 1e-5


def test_force_scales_inversely_with_a(oracle):
    d = _mesh_from_fn(16, lambda x: np.cos(2 * np.pi * x))
    e = oracle.poisson_accel(d, 16, 1.0, OMEGA_M, 0.5, 0.0)[0]
    l = oracle.poisson_accel(d, 16, 1.0, OMEGA_M, 1.0, 0.0)[0]
    assert abs(np.abs(e).max() / np.abs(l).max() - 2.0) < 1e-4


def test_nyquist_plane_has_no_gradient(oracle):
    # delta = cos(pi * ix) (the Nyquist mode along x): wavenumber() maps it to k = 0 and k^2 = 0,
    # so the reference drops the mode entirely (src/fft.rs:88-89, src/poisson.rs:63-65).
    n = 8
    d = np.broadcast_to(np.cos(np.pi * np.arange(n))[:, None, None], (n, n, n)).astype(np.float32).ravel()
    g = oracle.poisson_accel(d, n, 1.0, OMEGA_M, 1.0, 0.0, dtype=np.float64)
    assert np.max(np.abs(g)) < 1e-12


# ------------------------------------------------------------------ tree (src/tree.rs)

def test_short_range_factor_limits_and_golden(oracle):
    assert abs(oracle.short_range_factor(0.0, 1.0) - 1.0) < 1e-6
    assert oracle.short_range_factor(10.0, 1.0) < 1e-6
    prev = 1.0
    for s in range(1, 51):
        v = oracle.short_range_factor(s * 0.1, 1.0)
        assert v <= prev + 1e-7
        prev = v
    z = np.load(os.path.join(GOLD, "srfactor.npz"))
    s64 = np.array([oracle.short_range_factor(r, 1.0, np.float64) for r in z["r"]])
    s32 = np.array([oracle.short_range_factor(r, 1.0, np.float32) for r in z["r"]])
    assert np.max(np.abs(s64 - z["S"])) < 1e-14
    assert np.max(np.abs(s32 - z["S"])) < 3e-7


def test_bruteforce_matches_golden(oracle):
    z = np.load(os.path.join(GOLD, "bruteforce.npz"))
    for n, rc in ((400, 2.0), (3000, 2.5)):
        ref = z[f"acc{n}"]
        b64 = oracle.brute_force(z[f"pos{n}"], 10.0, 0.5, rc, 0.05, 1.0, dtype=np.float64)
        b32 = oracle.brute_force(z[f"pos{n}"], 10.0, 0.5, rc, 0.05, 1.0)
        assert max_diff(b64, ref) < 1e-12 * max_norm(ref)
        assert max_diff(b32, ref) < 1e-5 * max_norm(ref)


def test_tree_theta_zero_matches_brute_force(oracle):
    z = np.load(os.path.join(GOLD, "bruteforce.npz"))
    fast, _ = oracle.tree_accel(z["pos400"], 10.0, 0.5, 2.0, 0.05, 0.0, 1.0)
    assert max_diff(fast, z["acc400"]) / max_norm(z["acc400"]) < 1e-5


def test_tree_theta_half_close_to_brute_force(oracle):
    z = np.load(os.path.join(GOLD, "bruteforce.npz"))
    fast, _ = oracle.tree_accel(z["pos3000"], 10.0, 0.5, 2.5, 0.05, 0.5, 1.0)
    assert max_diff(fast, z["acc3000"]) / max_norm(z["acc3000"]) < 0.02
    # and the FP64 twin agrees with the FP32 oracle to FP32 round-off at theta = 0.5
    f64, _ = oracle.tree_accel(z["pos3000"], 10.0, 0.5, 2.5, 0.05, 0.5, 1.0, dtype=np.float64)
    assert max_diff(fast, f64) / max_norm(f64) < 1e-4


def test_periodic_pair_interacts_across_boundary(oracle):
    acc, _ = oracle.tree_accel([[0.1, 5.0, 5.0], [9.9, 5.0, 5.0]], 10.0, 0.5, 2.0, 0.05, 0.5, 1.0)
    assert acc[0, 0] < 0 and acc[1, 0] > 0


def test_inputs_match_reference_generators():
    # First values of the reference's xorshift streams, computed by hand from the recurrence.
    p = tree_scattered(1, 10.0, 7)
    s = 7
    M = (1 << 64) - 1
    s ^= (s << 13) & M
    s ^= s >> 7
    s ^= (s << 17) & M
    assert p[0, 0] == np.float32(np.float32(s >> 40) / np.float32(1 << 24)) * np.float32(10.0)


# ------------------------------------------------------- full TreePM (src/integrator.rs)

def test_treepm_pair_force_matches_newton(oracle):
    box, mesh_n = 100.0, 64
    split = oracle.force_split(box, mesh_n, 2, 0.05, 0.5)
    split["softening"] = 0.01
    acc, _, _ = oracle.accelerations([[50, 50, 50], [53, 50, 50]], mesh_n, box, OMEGA_M, 1.0, split)
    k = oracle.coupling(OMEGA_M, box, 1.0, 2)
    newton = k / 9.0
    assert acc[0, 0] > 0 and acc[1, 0] < 0
    assert abs(acc[0, 0] - newton) / newton < 0.05
    assert abs(acc[1, 0] + newton) / newton < 0.05


def test_fp32_oracle_close_to_fp64_twin(oracle):
    pos = tree_scattered(4096, 100.0, 5)
    split = oracle.force_split(100.0, 16, 16, 1 / 30, 0.5)
    a32, _, _ = oracle.accelerations(pos, 16, 100.0, OMEGA_M, 0.02, split)
    a64, _, _ = oracle.accelerations(pos.astype(np.float64), 16, 100.0, OMEGA_M, 0.02, split, dtype=np.float64)
    rel = np.sqrt(((a32 - a64) ** 2).sum(1)) / np.sqrt((a64 ** 2).sum(1))
    assert np.sqrt(np.mean(rel ** 2)) < 1e-4
