"""Per-step outputs on the GPU (src/runner.rs:262-284 snapshot, src/frame.rs:33-64) against the
oracle, bit for bit, on the reference's own frame test inputs and on a simulated state."""
import numpy as np
import pytest

from helpers import OMEGA_M

pytestmark = pytest.mark.gpu


def _ctx(treepm, n, box=100.0, cap=1 << 10):
    print(f"[gpu] frame context mesh={n}", flush=True)
    return treepm.Solver(n, box, OMEGA_M, max_particles=cap)


@pytest.mark.parametrize("n", [4, 48, 200, 256, 300])
def test_downsample_grid_matches_oracle(treepm, oracle, n):
    if n == 4:
        mesh = np.zeros(64, np.float32)
        mesh[(1 * 4 + 2) * 4 + 3] = 5.0  # src/frame.rs:70-81
    elif n == 200:
        mesh = np.zeros(n ** 3, np.float32)
        mesh[0] = 8.0  # src/frame.rs:83-93
    else:
        mesh = np.random.default_rng(n).normal(0, 2, n ** 3).astype(np.float32)
    s = _ctx(treepm, n)
    m_ref, ref = oracle.downsample_grid(mesh, n)
    m = s.downsample_grid_dev(0)
    assert m == m_ref
    d_in = treepm.DeviceArray.from_numpy(mesh)
    d_out = treepm.DeviceArray(m ** 3, np.float32)
    s.downsample_grid_dev(d_in.ptr, d_out.ptr)
    s.sync()
    got = d_out.numpy()
    assert np.array_equal(got, ref)
    if n == 4:
        assert got[(3 * 4 + 2) * 4 + 1] == 6.0
    if n == 200:
        assert abs(got[0] - 2.0) < 1e-6 and abs(got[1] - 1.0) < 1e-6


def test_snapshot_matches_oracle_on_a_simulated_state(treepm, oracle):
    n = 32
    box = 100.0
    split = treepm.force_split(box, n, n, 1 / 30, 0.5)
    s = treepm.Solver(n, box, OMEGA_M, split=split, max_particles=n ** 3)
    a0, a1 = 0.02, 0.05
    pos, mom = s.zeldovich(7, 2.0, 0.3 * box / n, treepm.zeldovich_momentum_factor(a0), want_momenta=True)
    pos, mom = s.run(pos, mom, a0, a1, 3)
    d_pos = treepm.DeviceArray.from_numpy(pos)
    d_mom = treepm.DeviceArray.from_numpy(mom)
    d_acc = treepm.DeviceArray(3 * n ** 3, np.float32)
    d_delta = treepm.DeviceArray(n ** 3, np.float32)
    s.accelerations_dev(d_pos.ptr, n ** 3, a1, d_acc.ptr, d_delta.ptr)
    outs = [treepm.DeviceArray(k * n ** 3, np.float32) for k in (3, 1, 1, 3)]
    _, _, e = treepm.growth(a1)
    s.snapshot_dev(d_pos.ptr, d_mom.ptr, n ** 3, a1, e, d_delta.ptr, *[o.ptr for o in outs])
    s.sync()
    got = [o.numpy() for o in outs]
    delta = d_delta.numpy()
    ref = oracle.snapshot(pos, mom, a1, e, delta, n, box)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r.reshape(-1))
    # densities are 1 + delta sampled at the particles: their mean is close to 1 (mass weighting
    # makes it slightly larger in a clustered field) and never negative
    dens = got[2]
    assert dens.min() >= 0.0 and 0.95 < dens.mean() < 1.5
    # only the outputs asked for are written
    only = treepm.DeviceArray(n ** 3, np.float32)
    s.snapshot_dev(d_pos.ptr, 0, n ** 3, a1, e, d_delta.ptr, 0, 0, only.ptr, 0)
    s.sync()
    assert np.array_equal(only.numpy(), ref[2])
    with pytest.raises(treepm.TpmError):
        s.snapshot_dev(d_pos.ptr, 0, n ** 3, a1, e, d_delta.ptr, 0, only.ptr, 0, 0)
