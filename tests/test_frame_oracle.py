"""Per-step outputs of the oracle (src/runner.rs:262-284 snapshot, src/frame.rs:33-64) pinned
by the reference's own frame tests and closed forms."""
import numpy as np

from helpers import OMEGA_M  # noqa: F401  (shared test constants)


def test_small_grid_keeps_every_cell_and_transposes(oracle):
    # src/frame.rs:70-81: n = 4, value 5 in mesh cell (1, 2, 3) -> texel x = 1, y = 2, z = 3 is 6.
    mesh = np.zeros(64, np.float32)
    mesh[(1 * 4 + 2) * 4 + 3] = 5.0
    m, g = oracle.downsample_grid(mesh, 4)
    assert m == 4
    assert g[(3 * 4 + 2) * 4 + 1] == 6.0
    assert np.count_nonzero(g == 1.0) == 63


def test_large_grid_averages_blocks(oracle):
    # src/frame.rs:83-93: n = 200 -> m = 100, delta = 8 in cell 0 -> first texel 2, the next 1.
    mesh = np.zeros(200 ** 3, np.float32)
    mesh[0] = 8.0
    m, g = oracle.downsample_grid(mesh, 200)
    assert m == 100
    assert abs(g[0] - 2.0) < 1e-6 and abs(g[1] - 1.0) < 1e-6


def test_downsample_block_means_against_numpy(oracle):
    # n = 300 -> f = 3, m = 100: every texel is 1 + the mean of its 3^3 block (numpy, FP64).
    rng = np.random.default_rng(5)
    n = 300
    mesh = rng.normal(0, 1, n ** 3).astype(np.float32)
    m, g = oracle.downsample_grid(mesh, n)
    assert m == 100
    ref = mesh.reshape(n, n, n).astype(np.float64).reshape(m, 3, m, 3, m, 3).mean(axis=(1, 3, 5)) + 1.0
    assert np.max(np.abs(g.reshape(m, m, m) - ref.transpose(2, 1, 0))) < 1e-5


def test_rsd_speed_and_density(oracle):
    # src/frame.rs:95-100: a = 0.5, E = 2, p = (1, 0, 0) -> shift 2; speed |p| / a = 2.
    pos = np.array([[5.0, 5.0, 5.0], [0.1, 9.9, 3.3]])
    mom = np.array([[1.0, 0.0, 0.0], [0.3, -0.4, 1.2]])
    n, box = 4, 10.0
    delta = np.zeros(n ** 3, np.float32)
    p32, speed, dens, rsd = oracle.snapshot(pos, mom, 0.5, 2.0, delta, n, box)
    assert np.allclose(rsd[0], [2.0, 0.0, 0.0]) and speed[0] == 2.0
    assert abs(speed[1] - 1.3 / 0.5) < 1e-6
    assert np.allclose(rsd[1], mom[1] / (0.25 * 2.0))
    assert np.array_equal(p32, pos.astype(np.float32))
    # uniform delta = 0 -> density 1; delta = -2 everywhere -> clamped at 0 (CIC never negative)
    assert np.all(dens == 1.0)
    _, _, dens2, _ = oracle.snapshot(pos, mom, 0.5, 2.0, np.full(n ** 3, -2.0, np.float32), n, box)
    assert np.all(dens2 == 0.0)


def test_density_matches_interpolation_of_a_linear_field(oracle):
    # CIC interpolation is exact for a field linear between cell centres: delta = x-index gives
    # 1 + (x / h - 0.5) away from the periodic seam.
    n, box = 16, 16.0
    ix = np.arange(n, dtype=np.float32)
    delta = np.broadcast_to(ix[:, None, None], (n, n, n)).astype(np.float32).ravel()
    rng = np.random.default_rng(3)
    pos = np.column_stack([rng.uniform(2, 13, 50), rng.uniform(0, 16, 50), rng.uniform(0, 16, 50)])
    mom = np.zeros_like(pos)
    _, _, dens, _ = oracle.snapshot(pos, mom, 1.0, 1.0, delta, n, box)
    assert np.max(np.abs(dens - (1.0 + pos[:, 0] - 0.5))) < 1e-5
