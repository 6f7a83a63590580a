"""GPU parity suite: libtpm's HIP path against the CPU oracle, the golden vectors and the
reference's own known-answer tests, at sizes the oracle finishes in seconds.

Tolerances (mass grid to 1e-6 relative, fixed-point grid bit-exact run to run; forces in FP32):
  deposit   |rho - rho_ref| <= 1e-6 max(rho_ref, mean rho) per cell; bit-exact run to run and
            under any permutation of the particles
  PM force  within 2e-5 of the field's maximum vs the FP64 oracle (FP32 FFT round-off)
  tree      theta = 0 within 1e-5 of max |a| of brute force (src/tree.rs:396-406)
"""
import os

import numpy as np
import pytest

from helpers import (OMEGA_M, fixed_point_delta32, lattice, max_diff, max_norm, mesh_scattered, rel_errors,
                     tree_scattered)

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def solver(treepm, n, box, split=None, max_particles=1 << 16, flags=0):
    print(f"[gpu] context mesh={n} box={box} split={split is not None} flags={flags} cap={max_particles}", flush=True)
    return treepm.Solver(n, box, OMEGA_M, split=split, max_particles=max_particles, flags=flags)


# ----------------------------------------------------------------------------- deposit

def test_deposit_kats(treepm):
    s = solver(treepm, 8, 1.0)
    h = 1.0 / 8
    g = s.deposit(np.array([[2.5 * h, 3.5 * h, 4.5 * h]], np.float32)).reshape(8, 8, 8)
    assert abs(g[2, 3, 4] - 1.0) < 1e-6 and (g > 1e-12).sum() == 1
    g = s.deposit(lattice(8, 1.0), 3.0)
    assert np.all(np.abs(g - 3.0) < 1e-6)
    s4 = solver(treepm, 4, 1.0)
    g = s4.deposit(np.array([[1.0, 0.125, 0.125]], np.float32)).reshape(4, 4, 4)
    assert abs(g[3, 0, 0] - 0.5) < 1e-6 and abs(g[0, 0, 0] - 0.5) < 1e-6


def test_deposit_matches_oracle_and_golden(treepm, oracle):
    z = np.load(os.path.join(GOLD, "cic_5000.npz"))
    s = solver(treepm, 16, 1.0)
    g32 = s.deposit(z["pos"], 0.25)
    g64 = s.deposit(z["pos"].astype(np.float64), 0.25)
    ref = z["grid"]
    scale = np.maximum(np.abs(ref), ref.mean())
    assert np.max(np.abs(g64 - ref) / scale) < 1e-6
    assert np.max(np.abs(g32 - ref) / scale) < 1e-6
    o32 = oracle.deposit(z["pos"], 16, 1.0, 0.25)
    assert np.max(np.abs(g32 - o32) / scale) < 1e-6
    assert abs(g32.sum(dtype=np.float64) - 0.25 * 5000) < 1e-6 * 0.25 * 5000


def test_deposit_bit_exact_and_order_independent(treepm):
    rng = np.random.default_rng(1)
    pos = (rng.random((50000, 3)) * 100.0).astype(np.float64)
    pos[:5000] = 50.0 + rng.standard_normal((5000, 3))  # a dense clump
    s = solver(treepm, 32, 100.0, max_particles=50000)
    g1 = s.deposit(pos)
    g2 = s.deposit(pos)
    g3 = s.deposit(pos[rng.permutation(len(pos))])
    assert np.array_equal(g1.view(np.uint32), g2.view(np.uint32))
    assert np.array_equal(g1.view(np.uint32), g3.view(np.uint32))


def test_dense_cube_deposit_in_evaluation(treepm, oracle):
    # The evaluation's deposit over Morton cubes takes the run-aggregated path for a cube of more
    # than one work item: 20,000 particles crowded into about one cell of a 32^3 mesh put far more
    # than 8,192 particles into one 16-cell cube. Its density contrast must match the oracle's (FP32
    # weights and FP64 weights), and the integer grid must not depend on the input order.
    rng = np.random.default_rng(5)
    box, n = 100.0, 32
    pos = rng.random((60000, 3)) * box
    pos[:20000] = 40.0 + rng.random((20000, 3)) * 3.0
    pos32 = pos.astype(np.float32)
    s = solver(treepm, n, box, max_particles=len(pos))
    _, d32 = s.accelerations(pos32, 0.5, want_delta=True)
    perm = rng.permutation(len(pos))
    _, d32p = s.accelerations(pos32[perm], 0.5, want_delta=True)
    assert np.array_equal(d32.view(np.uint32), d32p.view(np.uint32))
    # The fixed-point sum is exact, so the result is known bit for bit: the reference's FP32
    # weights, each particle's quanta summed as integers. (The FP32 oracle's serial FP32 sum over
    # the ~4,000 particles of a clump cell is itself 3e-6 off, and FP32 against FP64 weights
    # differ by 2.2e-6 of 1 + delta there.)
    assert np.array_equal(d32.view(np.uint32), fixed_point_delta32(pos32, n, box).view(np.uint32))
    ref32 = oracle.density_contrast(pos32, n, box)
    assert np.max(np.abs(d32 - ref32) / np.maximum(np.abs(ref32 + 1), 1.0)) < 5e-6
    _, d64 = s.accelerations(pos, 0.5, want_delta=True)
    ref64 = oracle.density_contrast(pos, n, box, dtype=np.float64)
    assert np.max(np.abs(d64 - ref64) / np.maximum(np.abs(ref64 + 1), 1.0)) < 1e-6
    assert d32.max() > 1000.0  # the clump really is dense


@pytest.mark.parametrize("n", [5, 12, 17, 20, 33, 48, 96])
def test_cube_deposit_exact_at_uneven_meshes(treepm, n):
    # Meshes whose side is no multiple of the deposit cube's (17, 33: cubes of 8.5 and 8.25 cells,
    # so the tiles take a ceil), meshes smaller than a tile (5, 12: one cube, tile cells wrapping
    # onto the same grid cells), and meshes of several whole cubes (20, 48, 96). At n = 12 the
    # single cube holds 20,000 particles: a dense cube of three work items. The evaluation's
    # density contrast must equal the exact integer sum of the reference's FP32 quanta bit for bit.
    rng = np.random.default_rng(100 + n)
    box = 100.0
    pos32 = (rng.random((20000, 3)) * box).astype(np.float32)
    s = solver(treepm, n, box, max_particles=len(pos32))
    _, d32 = s.accelerations(pos32, 0.5, want_delta=True)
    assert np.array_equal(d32.view(np.uint32), fixed_point_delta32(pos32, n, box).view(np.uint32))


def test_density_contrast_matches_oracle(treepm, oracle):
    pos = tree_scattered(20000, 100.0, 3)
    s = solver(treepm, 32, 100.0, max_particles=20000)
    d = s.density_contrast(pos)
    # FP32 interface: the CIC weights are the reference's FP32 ones, so compare with the FP32
    # oracle (only the summation differs: exact fixed point vs serial FP32).
    ref = oracle.density_contrast(pos, 32, 100.0)
    assert np.max(np.abs(d - ref) / np.maximum(np.abs(ref + 1), 1.0)) < 1e-6
    # FP64 interface: mass grid against the FP64 oracle deposit to 1e-6 relative.
    g = s.deposit(pos.astype(np.float64))
    g64 = oracle.deposit(pos.astype(np.float64), 32, 100.0, dtype=np.float64)
    assert np.max(np.abs(g - g64) / np.maximum(g64, g64.mean())) < 1e-6


# ------------------------------------------------------------------------- interpolate

def test_interpolate_matches_oracle(treepm, oracle):
    rng = np.random.default_rng(2)
    cells = rng.standard_normal(16 ** 3).astype(np.float32)
    pos = mesh_scattered(3000, 1.0)
    s = solver(treepm, 16, 1.0)
    v = s.interpolate(cells, pos)
    ref = oracle.interpolate(cells, 16, 1.0, pos)
    assert np.max(np.abs(v - ref)) <= 1e-6 * np.abs(cells).max()
    v = s.interpolate(np.full(16 ** 3, 2.5, np.float32), pos)
    assert np.all(np.abs(v - 2.5) < 1e-5)


# ---------------------------------------------------------------------------- PM force

def _mesh_from_fn(n, f):
    x = (np.arange(n) + 0.5) / n
    return np.broadcast_to(f(x)[:, None, None], (n, n, n)).astype(np.float32).ravel()


def test_plane_wave_force_matches_analytic(treepm):
    n, a = 16, 0.5
    k0 = 2 * np.pi
    c = 1.5 * OMEGA_M / a
    s = solver(treepm, n, 1.0)
    g = s.pm_accel(_mesh_from_fn(n, lambda x: np.cos(k0 * x)), a, 0.0)
    exp = _mesh_from_fn(n, lambda x: -(c / k0) * np.sin(k0 * x))
    amp = c / k0
    assert np.max(np.abs(g[0] - exp)) < 1e-5 * amp
    assert np.max(np.abs(g[1])) < 1e-5 * amp and np.max(np.abs(g[2])) < 1e-5 * amp


def test_uniform_delta_and_inverse_a(treepm):
    s = solver(treepm, 8, 1.0)
    assert np.max(np.abs(s.pm_accel(np.ones(512, np.float32), 1.0))) < 1e-5
    s16 = solver(treepm, 16, 1.0)
    d = _mesh_from_fn(16, lambda x: np.cos(2 * np.pi * x))
    r = np.abs(s16.pm_accel(d, 0.5)[0]).max() / np.abs(s16.pm_accel(d, 1.0)[0]).max()
    assert abs(r - 2.0) < 1e-4


def test_wavefront_transforms_at_512(treepm):
    # The n = 512 wavefront path: plane waves along each axis and an oblique one against the
    # closed form g = -(1.5 Omega_m / a) k sin(k.x) / k^2 (src/poisson.rs:114-139), then a random
    # field against the LDS line kernels on the same context.
    n, a = 512, 0.5
    c = 1.5 * OMEGA_M / a
    print("[gpu] PM at 512^3, wavefront transforms", flush=True)
    s = solver(treepm, n, 1.0)
    x = (np.arange(n) + 0.5) / n
    for kv in ((7, 0, 0), (0, 5, 0), (0, 0, 9), (3, -5, 7)):
        k = 2 * np.pi * np.array(kv, np.float64)
        phase = (k[0] * x[:, None, None] + k[1] * x[None, :, None] + k[2] * x[None, None, :])
        d = np.cos(phase).astype(np.float32).ravel()
        g = s.pm_accel(d, a, 0.0)
        sn = np.sin(phase).ravel()
        k2 = (k ** 2).sum()
        amp = c / np.sqrt(k2)
        for ax in range(3):
            exp = -c * k[ax] * sn / k2
            assert np.max(np.abs(g[ax] - exp)) < 2e-5 * amp, (kv, ax)
        del g, sn, d, phase
    rng = np.random.default_rng(512)
    d = rng.standard_normal(n ** 3).astype(np.float32)
    rs = 1.25 / n
    g_wave = s.pm_accel(d, 0.3, rs)
    s.set_tuning(16, 0, fft_path=1)
    g_line = s.pm_accel(d, 0.3, rs)
    assert np.max(np.abs(g_wave - g_line)) < 2e-5 * np.abs(g_line).max()


@pytest.mark.parametrize("n", [4, 6, 8, 14, 16, 24, 30, 32, 47, 64])
def test_pm_random_delta_matches_oracle(treepm, oracle, n):
    rng = np.random.default_rng(n)
    d = rng.standard_normal(n ** 3).astype(np.float32)
    s = solver(treepm, n, 100.0)
    rs = 1.25 * 100.0 / n
    g = s.pm_accel(d, 0.3, rs)
    ref = oracle.poisson_accel(d, n, 100.0, OMEGA_M, 0.3, rs, dtype=np.float64)
    assert np.max(np.abs(g - ref)) < 2e-5 * np.abs(ref).max()


# -------------------------------------------------------------------------- short range

def test_tree_theta_zero_matches_brute_force(treepm):
    z = np.load(os.path.join(GOLD, "bruteforce.npz"))
    split = {"r_split": 0.5, "r_cut": 2.0, "softening": 0.05, "theta": 0.0}
    for flags in (0, treepm.FLAG_REF_WALK):
        s = solver(treepm, 8, 10.0, split, flags=flags)
        fast = s.short_range(z["pos400"], 1.0)
        assert max_diff(fast, z["acc400"]) / max_norm(z["acc400"]) < 1e-5


def test_tree_theta_half_close_to_brute_force(treepm, oracle):
    z = np.load(os.path.join(GOLD, "bruteforce.npz"))
    split = {"r_split": 0.5, "r_cut": 2.5, "softening": 0.05, "theta": 0.5}
    s = solver(treepm, 8, 10.0, split)
    fast = s.short_range(z["pos3000"], 1.0)
    assert max_diff(fast, z["acc3000"]) / max_norm(z["acc3000"]) < 0.02
    # reference walk (same octree rule as the reference) against the oracle's tree
    for flags in (treepm.FLAG_REF_WALK, treepm.FLAG_REF_WALK | treepm.FLAG_REF_LEAF_RULE):
        s = solver(treepm, 8, 10.0, split, flags=flags)
        ref_gpu = s.short_range(z["pos3000"], 1.0)
        ora, _ = oracle.tree_accel(z["pos3000"], 10.0, 0.5, 2.5, 0.05, 0.5, 1.0)
        assert max_diff(ref_gpu, z["acc3000"]) / max_norm(z["acc3000"]) < 0.02
        if flags & treepm.FLAG_REF_LEAF_RULE:
            assert max_diff(ref_gpu, ora) / max_norm(ora) < 1e-4


def test_periodic_pair_interacts_across_boundary(treepm):
    split = {"r_split": 0.5, "r_cut": 2.0, "softening": 0.05, "theta": 0.5}
    s = solver(treepm, 8, 10.0, split)
    acc = s.short_range(np.array([[0.1, 5.0, 5.0], [9.9, 5.0, 5.0]], np.float32), 1.0)
    assert acc[0, 0] < 0 and acc[1, 0] > 0
    assert abs(acc[0, 0] + acc[1, 0]) < 1e-6 * abs(acc[0, 0])


# ---------------------------------------------------------------------- full TreePM

def test_treepm_pair_force_matches_newton(treepm):
    split = treepm.force_split(100.0, 64, 2, 0.05, 0.5)
    split["softening"] = 0.01
    s = solver(treepm, 64, 100.0, split)
    acc, _ = s.accelerations(np.array([[50, 50, 50], [53, 50, 50]], np.float32), 1.0)
    newton = treepm.coupling(OMEGA_M, 100.0, 1.0, 2) / 9.0
    assert acc[0, 0] > 0 and acc[1, 0] < 0
    assert abs(acc[0, 0] - newton) / newton < 0.05 and abs(acc[1, 0] + newton) / newton < 0.05


def _zeldovich_config(treepm, n, seed=1):
    s = solver(treepm, n, 100.0, max_particles=n ** 3)
    pos, _ = s.zeldovich(seed, 2.0, 0.3 * 100.0 / n)
    s.close()
    return pos


def test_zeldovich_matches_numpy_twin(treepm):
    import ic_twin
    n = 16
    s = solver(treepm, n, 100.0, max_particles=n ** 3)
    pos, mom = s.zeldovich(7, 2.0, 0.3 * 100.0 / n, mom_factor=0.5, want_momenta=True)
    ref, rmom = ic_twin.zeldovich(n, 100.0, 7, 2.0, 0.3 * 100.0 / n, mom_factor=0.5)
    d = pos - ref
    d -= 100.0 * np.round(d / 100.0)
    assert np.max(np.abs(d)) < 1e-9 * 100.0
    assert np.max(np.abs(mom - rmom)) < 1e-9 * np.abs(rmom).max()
    disp = mom / 0.5
    assert abs(np.sqrt((disp ** 2).sum(1).mean()) - 0.3 * 100.0 / n) < 1e-9


def test_bbks_zeldovich_matches_numpy_twin(treepm):
    # The reference's default ICs (src/runner.rs:194-206): sigma_8-normalised BBKS spectrum at z = 500,
    # x = q + D Psi, p = a^2 f E D Psi; the same noise through numpy's FFT.
    import ic_twin
    n, box = 16, 100.0
    a0 = 1.0 / 501.0
    s = solver(treepm, n, box, max_particles=n ** 3)
    pos, mom = s.zeldovich_bbks(5, a0)
    _, amp = treepm.power_spectrum(0.1)
    gamma = treepm.Cosmology.lcdm_default().omega_m * treepm.H0 / 100.0
    d, f, e = treepm.growth(a0)
    ref, rmom = ic_twin.zeldovich_bbks(n, box, 5, amp, treepm.N_S, gamma, d, a0 * a0 * f * e)
    dd = pos - ref
    dd -= box * np.round(dd / box)
    assert np.max(np.abs(dd)) < 1e-9 * box
    assert np.max(np.abs(mom - rmom)) < 1e-9 * np.abs(rmom).max()
    # the displacement at z = 500 is a small fraction of a cell (D ~ 0.0026)
    disp = np.sqrt((mom / (a0 * a0 * f * e) ** 1) ** 2).max()
    assert 0 < disp < 0.5 * box / n


@pytest.mark.parametrize("n", [16, 32])
def test_full_treepm_matches_oracle(treepm, oracle, n):
    pos = _zeldovich_config(treepm, n).astype(np.float32)
    split = treepm.force_split(100.0, n, n, 1 / 30, 0.5)
    a = 1.0 / 50.0
    s = solver(treepm, n, 100.0, split, max_particles=n ** 3)
    acc, delta = s.accelerations(pos, a, want_delta=True)
    ref, dref, _ = oracle.accelerations(pos, n, 100.0, OMEGA_M, a, split, want_delta=True)
    exact, _, _ = oracle.accelerations(pos.astype(np.float64), n, 100.0, OMEGA_M, a,
                                       dict(split, theta=0.0), dtype=np.float64)
    assert np.max(np.abs(delta - dref)) < 1e-5 * max(1.0, np.abs(dref).max())
    e_gpu = rel_errors(acc, exact)
    e_ora = rel_errors(ref, exact)
    print(f"[gpu] n={n} rms rel err gpu={np.sqrt(np.mean(e_gpu**2)):.3e} oracle={np.sqrt(np.mean(e_ora**2)):.3e}")
    # The group walk opens every node the reference opens, so it is at least as accurate.
    assert np.sqrt(np.mean(e_gpu ** 2)) <= np.sqrt(np.mean(e_ora ** 2)) * 1.05 + 1e-5
    # The reference-walk mode reproduces the oracle to FP32 round-off.
    s2 = solver(treepm, n, 100.0, split, max_particles=n ** 3,
                flags=treepm.FLAG_REF_WALK | treepm.FLAG_REF_LEAF_RULE)
    acc2, _ = s2.accelerations(pos, a)
    assert np.sqrt(np.mean(rel_errors(acc2, ref) ** 2)) < 1e-4


def test_force_evaluation_is_bit_exact_run_to_run(treepm):
    n = 32
    pos = _zeldovich_config(treepm, n)
    split = treepm.force_split(100.0, n, n, 1 / 30, 0.5)
    s = solver(treepm, n, 100.0, split, max_particles=n ** 3)
    a1, d1 = s.accelerations(pos, 0.02, want_delta=True)
    a2, d2 = s.accelerations(pos, 0.02, want_delta=True)
    assert np.array_equal(a1.view(np.uint32), a2.view(np.uint32))
    assert np.array_equal(d1.view(np.uint32), d2.view(np.uint32))


def test_second_stream_does_not_change_the_forces(treepm):
    # Above 2^22 particles the octree and walk groups run on a second stream beside the deposit,
    # FFT and interpolation; with that switched off every stage runs on one stream in turn. The
    # forces must be the same bit for bit, and so must a second evaluation with the overlap back on.
    n = 192  # 7.1M particles
    npart = n ** 3
    split = treepm.force_split(100.0, n, n, 1 / 30, 0.35)
    s = solver(treepm, n, 100.0, split, max_particles=npart)
    d_pos = treepm.DeviceArray(3 * npart, np.float64)
    d_acc = treepm.DeviceArray(3 * npart, np.float32)
    s.zeldovich_dev(1, 2.0, 0.3 * 100.0 / n, 0.0, d_pos.ptr)
    s.accelerations_dev(d_pos.ptr, npart, 0.02, d_acc.ptr)
    s.sync()
    beside = d_acc.numpy()
    s.set_overlap(False)
    s.accelerations_dev(d_pos.ptr, npart, 0.02, d_acc.ptr)
    s.sync()
    alone = d_acc.numpy()
    s.set_overlap(True)
    s.accelerations_dev(d_pos.ptr, npart, 0.02, d_acc.ptr)
    s.sync()
    again = d_acc.numpy()
    assert np.all(np.isfinite(beside)) and np.abs(beside).max() > 0
    assert np.array_equal(beside.view(np.uint32), alone.view(np.uint32))
    assert np.array_equal(beside.view(np.uint32), again.view(np.uint32))


def test_group_building_paths_agree(treepm):
    # Walk groups come from per-cube starts when the cube count fits the context's scratch and from
    # scans over all particles otherwise; both give the same groups, hence bit-identical forces.
    # 16^3 particles with the group level raised by 2: 8^4 = 4096 cubes, so a context for exactly
    # 4096 particles takes the scan path and one for 8192 the per-cube path.
    n = 16
    pos = _zeldovich_config(treepm, n)
    split = treepm.force_split(100.0, n, n, 1 / 30, 0.5)
    out = []
    for cap in (n ** 3, 2 * n ** 3):
        s = solver(treepm, n, 100.0, split, max_particles=cap)
        s.set_tuning(16, 2, fft_path=0, tree_depth=21, small_node_direct=0)
        s.set_timing(True)
        a, _ = s.accelerations(pos, 0.02)
        out.append((a, s.stats()["groups"]))
        s.close()
    assert out[0][1] == out[1][1] > 0
    assert np.array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))


def test_depth_limit_is_raised_to_the_group_cubes(treepm):
    # A depth limit of 8 with the group level at 9 (level 2 for 4096 particles, shifted by 7): the
    # sort is taken to level 9, so each group cube's particles stay contiguous. 4096 particles in
    # one level-8 cube: sorted only to level 8 they would stay in their random input order and cut
    # the eight level-9 cubes into thousands of runs; sorted to level 9 each cube of c particles
    # gives ceil(c / 64) groups, as with no depth limit. At theta = 0 every pair is summed, so the
    # forces agree to FP32 summation order.
    rng = np.random.default_rng(5)
    pos = (25 + rng.random((4096, 3))) * (100.0 / 256)
    cells = np.floor(pos / (100.0 / 512)).astype(np.int64)
    _, counts = np.unique(cells, axis=0, return_counts=True)
    assert len(counts) == 8
    groups = int(np.sum((counts + 63) // 64))
    split = treepm.force_split(100.0, 16, 16, 1 / 30, 0.0)
    out = []
    for depth in (8, 21):
        s = solver(treepm, 16, 100.0, split, max_particles=4096)
        s.set_tuning(16, 7, fft_path=0, tree_depth=depth, small_node_direct=0)
        s.set_timing(True)
        a, _ = s.accelerations(pos, 0.02)
        out.append((a.astype(np.float64), s.stats()["groups"]))
        s.close()
    assert out[0][1] == out[1][1] == groups
    assert max_diff(out[0][0], out[1][0]) / max_norm(out[1][0]) < 1e-5


def test_walk_without_counting_gives_identical_forces(treepm):
    # The interaction counter is instrumentation only: switching it off changes no bit of the
    # forces; the interaction total then reads 0 while the other walk statistics stay.
    n = 32
    pos = _zeldovich_config(treepm, n)
    split = treepm.force_split(100.0, n, n, 1 / 30, 0.5)
    s = solver(treepm, n, 100.0, split, max_particles=n ** 3)
    s.set_timing(True)
    a1, _ = s.accelerations(pos, 0.02)
    st1 = s.stats()
    s.set_counting(False)
    a2, _ = s.accelerations(pos, 0.02)
    st2 = s.stats()
    s.set_counting(True)
    a3, _ = s.accelerations(pos, 0.02)
    st3 = s.stats()
    assert np.array_equal(a1.view(np.uint32), a2.view(np.uint32))
    assert np.array_equal(a1.view(np.uint32), a3.view(np.uint32))
    assert st1["interactions"] > 0 and st2["interactions"] == 0 and st3["interactions"] == st1["interactions"]
    assert st2["kept"] == st1["kept"] and st2["listed"] == st1["listed"]


def test_judge_matches_fp64_oracle(treepm, oracle):
    n = 16
    pos = _zeldovich_config(treepm, n)
    split = treepm.force_split(100.0, n, n, 1 / 30, 0.5)
    a = 0.02
    idx = np.arange(0, n ** 3, 7, dtype=np.uint32)
    s = solver(treepm, n, 100.0, split, max_particles=n ** 3, flags=treepm.FLAG_REF_LEAF_RULE)
    j = s.judge(pos, idx, a, theta=0.2)
    ref, _, _ = oracle.accelerations(pos, n, 100.0, OMEGA_M, a, dict(split, theta=0.2), dtype=np.float64)
    assert np.max(rel_errors(j, ref[idx])) < 1e-8


def test_judge_does_not_depend_on_tuning(treepm):
    # The judge builds the full-depth octree whatever the tuned depth limit, so its output is the
    # same before and after tuning the product's octree to depth 8.
    n = 32
    pos = _zeldovich_config(treepm, n)
    split = treepm.force_split(100.0, n, n, 1 / 30, 0.5)
    idx = np.arange(0, n ** 3, 5, dtype=np.uint32)
    s = solver(treepm, n, 100.0, split, max_particles=n ** 3)
    j1 = s.judge(pos, idx, 0.02, theta=0.2)
    s.set_tuning(16, 0, fft_path=0, tree_depth=8, small_node_direct=0)
    s.accelerations(pos, 0.02)
    j2 = s.judge(pos, idx, 0.02, theta=0.2)
    assert np.array_equal(j1, j2)


def test_edge_cases(treepm):
    split = treepm.force_split(100.0, 16, 16, 1 / 30, 0.5)
    s = solver(treepm, 16, 100.0, split, max_particles=1000, flags=treepm.FLAG_CHECK_FINITE)
    acc, _ = s.accelerations(np.zeros((0, 3), np.float32), 0.5)
    assert acc.shape == (0, 3)
    acc, _ = s.accelerations(np.array([[10.0, 20.0, 30.0]], np.float32), 0.5)
    # One particle feels no force from its own density: CIC deposit and interpolation with an
    # odd gradient kernel cancel exactly, up to FP32 round-off of the 4096x overdensity.
    newton_one_cell = treepm.coupling(OMEGA_M, 100.0, 0.5, 1) / (100.0 / 16) ** 2
    assert np.all(np.abs(acc) < 1e-4 * newton_one_cell)
    # positions exactly on the box face and ragged counts
    rng = np.random.default_rng(5)
    pos = (rng.random((999, 3)) * 100).astype(np.float32)
    pos[:3] = 100.0
    acc, _ = s.accelerations(pos, 0.5)
    assert np.all(np.isfinite(acc))
    with pytest.raises(treepm.TpmError):
        s.accelerations(np.zeros((1001, 3), np.float32), 0.5)
