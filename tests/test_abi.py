"""The C-ABI library loads and exports every entry point include/tpm.h declares; argument
errors come back as return codes (no GPU work is started by these tests)."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(REPO, "include", "tpm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tpm_[a-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree(treepm):
    assert header_functions() == sorted(treepm.EXPORTS)


def test_library_exports_every_declared_symbol(treepm):
    lib = ctypes.CDLL(treepm.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_version_string(treepm):
    assert b"gfx950" in treepm.load().tpm_version()


def test_invalid_params_are_rejected_without_device_work(treepm):
    lib = treepm.load()
    p = treepm.Params()
    p.mesh_n = 1  # below the minimum
    p.box_size = 100.0
    h = ctypes.c_void_p()
    rc = lib.tpm_create(ctypes.byref(p), 16, ctypes.byref(h))
    assert rc == treepm.TPM_E_INVALID
    assert b"mesh_n" in lib.tpm_last_error(None)
    p.mesh_n = 16
    p.flags = treepm.FLAG_TREEPM  # TreePM without r_split
    rc = lib.tpm_create(ctypes.byref(p), 16, ctypes.byref(h))
    assert rc == treepm.TPM_E_INVALID
    assert b"r_split" in lib.tpm_last_error(None)
    assert lib.tpm_create(None, 16, ctypes.byref(h)) == treepm.TPM_E_INVALID


def test_null_context_calls_fail_cleanly(treepm):
    lib = treepm.load()
    assert lib.tpm_accelerations_f32(None, None, 0, 1.0, None, None) == treepm.TPM_E_INVALID
    assert lib.tpm_set_timing(None, 1) == treepm.TPM_E_INVALID
    lib.tpm_destroy(None)


def test_force_split_matches_reference_formula(treepm):
    # ForceSplit::with (src/tree.rs:40-50) for L = 100, mesh 64, 64 particles/side, eps = spacing/30
    s = treepm.force_split(100.0, 64, 64, 1 / 30, 0.5)
    assert s["r_split"] == pytest.approx(1.953125)
    assert s["r_cut"] == pytest.approx(8.7890625)
    assert s["softening"] == pytest.approx(100 / 64 / 30, rel=1e-6)
