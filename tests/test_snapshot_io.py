"""Binary snapshot files (the reference plans them, plan.md:160-163: positions, velocities, a, box
size, parameters; it defines no format): the header layout of tpm.h, round trips, and the errors
for files that are not snapshots. Host code only, no GPU."""
import ctypes
import os

import numpy as np
import pytest


def _header(treepm, **kw):
    h = treepm.SnapshotHeader()
    h.a, h.box_size, h.omega_m, h.omega_l, h.omega_r, h.h0 = 0.25, 100.0, 0.315, 0.6849086, 9.14e-5, 67.4
    h.softening, h.r_split, h.r_cut, h.theta = 0.125, 3.90625, 17.578125, 0.5
    h.mesh_n, h.step, h.seed = 80, 12, 7
    for k, v in kw.items():
        setattr(h, k, v)
    return h


def test_header_is_128_bytes(treepm):
    assert ctypes.sizeof(treepm.SnapshotHeader) == 128


def test_round_trip_and_layout(treepm, tmp_path):
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 100, (1000, 3))
    mom = rng.normal(0, 1, (1000, 3))
    path = tmp_path / "s.tpmsnap"
    treepm.snapshot_write(path, _header(treepm), pos, mom)
    raw = path.read_bytes()
    # layout: 128-byte header, positions, momenta (little-endian doubles)
    assert len(raw) == 128 + 2 * 3 * 1000 * 8
    assert raw[:8] == b"TPMSNAP1"
    assert int.from_bytes(raw[8:12], "little") == 1 and int.from_bytes(raw[12:16], "little") == 128
    assert int.from_bytes(raw[16:24], "little") == 1000
    assert np.array_equal(np.frombuffer(raw[128:128 + 24000], "<f8").reshape(-1, 3), pos)
    assert np.array_equal(np.frombuffer(raw[128 + 24000:], "<f8").reshape(-1, 3), mom)
    h, p2, m2 = treepm.snapshot_read(path)
    assert np.array_equal(p2, pos) and np.array_equal(m2, mom)
    ref = _header(treepm).as_dict()
    got = h.as_dict()
    for k in ("a", "box_size", "omega_m", "omega_l", "omega_r", "h0", "softening", "r_split", "r_cut", "theta",
              "mesh_n", "step", "seed"):
        assert got[k] == ref[k], k
    assert got["n"] == 1000 and got["magic"] == b"TPMSNAP1"
    # header only
    assert treepm.snapshot_read(path, arrays=False).n == 1000


def test_empty_snapshot(treepm, tmp_path):
    path = tmp_path / "empty.tpmsnap"
    treepm.snapshot_write(path, _header(treepm), np.zeros((0, 3)), np.zeros((0, 3)))
    h, p, m = treepm.snapshot_read(path)
    assert h.n == 0 and p.shape == (0, 3) and m.shape == (0, 3)


def test_rejects_other_files_and_small_buffers(treepm, tmp_path):
    bad = tmp_path / "bad.bin"
    bad.write_bytes(os.urandom(512))
    with pytest.raises(treepm.TpmError) as e:
        treepm.snapshot_read(bad)
    assert e.value.code == treepm.TPM_E_IO
    with pytest.raises(treepm.TpmError) as e:
        treepm.snapshot_read(tmp_path / "missing.tpmsnap")
    assert e.value.code == treepm.TPM_E_IO
    # truncated arrays
    path = tmp_path / "t.tpmsnap"
    treepm.snapshot_write(path, _header(treepm), np.ones((10, 3)), np.ones((10, 3)))
    path.write_bytes(path.read_bytes()[:128 + 100])
    with pytest.raises(treepm.TpmError):
        treepm.snapshot_read(path)
    # capacity below n
    treepm.snapshot_write(path, _header(treepm), np.ones((10, 3)), np.ones((10, 3)))
    h = treepm.SnapshotHeader()
    buf = np.empty((5, 3))
    rc = treepm.load().tpm_snapshot_read(os.fsencode(path), ctypes.byref(h), buf.ctypes.data_as(ctypes.c_void_p),
                                         None, 5)
    assert rc == treepm.TPM_E_INVALID
    with pytest.raises(ValueError):
        treepm.snapshot_write(path, _header(treepm), np.ones((10, 3)), np.ones((9, 3)))
