"""The allocation-free deposit (tpm_deposit_dev_f64_noalloc) against tpm_deposit_f64.

Both are the fixed-point CIC deposit (src/mesh.rs:62-69 with 2^-32 quanta), so the grids must agree
bit for bit. Two calls in a row share the library's built-in grid; the second, smaller one checks
that the cells the previous call left are cleared. Device buffers come from libtpm's own runtime.
"""
import ctypes

import numpy as np
import pytest

from helpers import OMEGA_M

pytestmark = pytest.mark.gpu

BOX = 100.0


def _positions(count, seed):
    rng = np.random.default_rng(seed)
    pos = rng.random((count, 3)) * BOX
    pos[: count // 10] = 50.0 + 2.0 * rng.standard_normal((count // 10, 3))  # a dense clump
    pos[:3] = [0.0, 0.0, 0.0]  # on the box corner: the stencil wraps
    return np.mod(pos, BOX)


def _noalloc(treepm, pos, mesh, mass):
    lib = treepm.load()
    n = len(pos)
    d_pos = treepm.DeviceArray(max(3 * n, 1), np.float64)
    if n:
        d_pos.upload(pos.ravel())
    d_cells = treepm.DeviceArray(mesh ** 3, np.float32)
    print(f"[gpu] noalloc deposit n={n} mesh={mesh}", flush=True)
    rc = lib.tpm_deposit_dev_f64_noalloc(ctypes.c_void_p(d_pos.ptr), n, mesh, BOX, mass, ctypes.c_void_p(d_cells.ptr),
                                         None)
    assert rc == treepm.TPM_OK, lib.tpm_last_error(None)
    out = d_cells.numpy()  # a blocking copy on the null stream, after the deposit
    d_pos.free()
    d_cells.free()
    return out


def _with_context(treepm, pos, mesh, mass):
    with treepm.Solver(mesh, BOX, OMEGA_M, max_particles=max(len(pos), 1)) as s:
        return s.deposit(pos, mass)


def test_noalloc_deposit_matches_context_deposit_twice(treepm):
    pos1 = _positions(300_000, 1)
    pos2 = _positions(20_000, 2)
    first = _noalloc(treepm, pos1, 96, 1.0)
    second = _noalloc(treepm, pos2, 32, 0.5)
    ref1 = _with_context(treepm, pos1, 96, 1.0)
    ref2 = _with_context(treepm, pos2, 32, 0.5)
    assert np.array_equal(first.view(np.uint32), ref1.view(np.uint32))
    assert np.array_equal(second.view(np.uint32), ref2.view(np.uint32))
    assert abs(second.sum(dtype=np.float64) - 0.5 * len(pos2)) < 1e-6 * len(pos2)


def test_noalloc_deposit_largest_mesh_then_empty(treepm):
    pos = _positions(100_000, 3)
    full = _noalloc(treepm, pos, 128, 1.0)
    assert np.array_equal(full.view(np.uint32), _with_context(treepm, pos, 128, 1.0).view(np.uint32))
    # no particles: every cell the previous call filled must read zero
    empty = _noalloc(treepm, np.zeros((0, 3)), 128, 1.0)
    assert not np.any(empty)


def test_noalloc_deposit_rejects_oversized_mesh(treepm):
    lib = treepm.load()
    cells = treepm.DeviceArray(8, np.float32)
    pos = treepm.DeviceArray(3, np.float64)
    rc = lib.tpm_deposit_dev_f64_noalloc(ctypes.c_void_p(pos.ptr), 1, 129, BOX, 1.0, ctypes.c_void_p(cells.ptr), None)
    assert rc == treepm.TPM_E_INVALID
    assert b"128" in lib.tpm_last_error(None)
    assert lib.tpm_deposit_dev_f64_noalloc(ctypes.c_void_p(pos.ptr), 1, 16, BOX, 1.0, None, None) == treepm.TPM_E_INVALID
    pos.free()
    cells.free()
