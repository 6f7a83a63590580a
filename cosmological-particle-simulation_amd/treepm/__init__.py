"""Python binding of libtpm (the C ABI in include/tpm.h), used by the tests and the benchmark.

It mirrors the reference's interface for the force path (src/integrator.rs:140-161 and the
public functions under it) so that the tests read like the reference's own:

    Solver(params, max_particles)            PoissonSolver::new + the tree (src/poisson.rs:23)
    Solver.accelerations(pos, a)             Simulation::accelerations (src/integrator.rs:140)
    Solver.deposit(pos, mass)                Mesh::deposit (src/mesh.rs:62)
    Solver.density_contrast(pos)             Simulation::density_contrast (src/integrator.rs:117)
    Solver.pm_accel(delta, a, r_split)       PoissonSolver::acceleration (src/poisson.rs:33)
    Solver.interpolate(cells, pos)           Mesh::interpolate (src/mesh.rs:72)
    Solver.short_range(pos, k)               tree::accelerations (src/tree.rs:314)
    force_split(...)                         ForceSplit::with (src/tree.rs:40)

Every call goes through the HIP library; there is no CPU fallback. Loading fails loudly when
libtpm.so has not been built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libtpm.so")

TPM_OK = 0
TPM_E_INVALID = -1
TPM_E_DEVICE = -2
TPM_E_CAPACITY = -3
TPM_E_NONFINITE = -4
TPM_E_IO = -5

FLAG_TREEPM = 1
FLAG_REF_LEAF_RULE = 2
FLAG_CHECK_FINITE = 4
FLAG_REF_WALK = 8

STATUS_WALK_OVERFLOW = 1
STATUS_TREE_CAPACITY = 2
STATUS_NONFINITE = 4

# Symbols the header declares (checked by the CPU test suite against the built library).
EXPORTS = [
    "tpm_version", "tpm_create", "tpm_destroy", "tpm_last_error", "tpm_set_params",
    "tpm_accelerations_f32", "tpm_accelerations_f64", "tpm_accelerations_dev_f64",
    "tpm_get_stats", "tpm_set_timing", "tpm_set_counting", "tpm_set_overlap", "tpm_deposit_f32",
    "tpm_deposit_f64",
    "tpm_density_contrast_f32", "tpm_pm_accel", "tpm_interpolate_f32",
    "tpm_short_range_f32", "tpm_short_range_f64", "tpm_judge_f64",
    "tpm_zeldovich", "tpm_zeldovich_dev", "tpm_kdk_factors", "tpm_step_to_dev", "tpm_run",
    "tpm_device_alloc", "tpm_device_free", "tpm_memcpy", "tpm_sync", "tpm_nfw_halos", "tpm_nfw_halos_dev",
    "tpm_mlp_param_count", "tpm_mlp_init", "tpm_mlp_get_params", "tpm_mlp_set_params", "tpm_mlp_gradients",
    "tpm_mlp_predict", "tpm_mlp_train_dev", "tpm_accelerations_learned_dev",
    "tpm_growth", "tpm_power_spectrum", "tpm_power_sigma", "tpm_zeldovich_bbks", "tpm_zeldovich_bbks_dev",
    "tpm_snapshot_dev", "tpm_run_settings_default", "tpm_run_settings_sanitized", "tpm_headless_run", "tpm_downsample_grid_dev", "tpm_autotune", "tpm_set_tuning", "tpm_ensemble_create", "tpm_ensemble_destroy", "tpm_ensemble_last_error",
    "tpm_ensemble_ics", "tpm_ensemble_autotune", "tpm_ensemble_run", "tpm_ensemble_get",
    "tpm_snapshot_write", "tpm_snapshot_read", "tpm_deposit_dev_f64_noalloc", "tpm_get_device_status",
    "tpm_morton_order", "tpm_ensemble_graph_nodes", "tpm_mlp_error_scale",
]


# The reference's spectrum constants (src/ics.rs:14-18) and Hubble constant (src/cosmology.rs:27).
N_S = 0.965
SIGMA8 = 0.811
H0 = 67.4


class TpmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tpm error {code}: {msg}")
        self.code = code


class Cosmology(ctypes.Structure):
    _fields_ = [("omega_r", ctypes.c_double), ("omega_m", ctypes.c_double), ("omega_l", ctypes.c_double)]

    @classmethod
    def lcdm_default(cls):
        """Cosmology::lcdm_default (src/cosmology.rs:21-30), the float32 values of the reference."""
        f = np.float32
        c = cls()
        c.omega_r = float(f(9.14e-5))
        c.omega_m = float(f(0.315))
        c.omega_l = float(f(1.0) - f(0.315) - f(9.14e-5))
        return c


class NfwSpec(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("n_halos", ctypes.c_int32), ("halo_fraction", ctypes.c_double),
                ("mass_min", ctypes.c_double), ("mass_max", ctypes.c_double), ("slope", ctypes.c_double),
                ("conc_min", ctypes.c_double), ("conc_max", ctypes.c_double)]

    @classmethod
    def clustered(cls, seed=1, n_halos=20000, halo_fraction=0.7):
        """The clustered stress workload: 20,000 halos, M^-1.9 on [1e11, 1e15], c in [5, 15], 30% background."""
        s = cls()
        s.seed, s.n_halos, s.halo_fraction = int(seed), int(n_halos), float(halo_fraction)
        s.mass_min, s.mass_max, s.slope = 1e11, 1e15, 1.9
        s.conc_min, s.conc_max = 5.0, 15.0
        return s


class Params(ctypes.Structure):
    _fields_ = [
        ("mesh_n", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("box_size", ctypes.c_double),
        ("omega_m", ctypes.c_double),
        ("r_split", ctypes.c_double),
        ("r_cut", ctypes.c_double),
        ("softening", ctypes.c_double),
        ("theta", ctypes.c_double),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("ms_total", ctypes.c_double),
        ("ms_sort", ctypes.c_double),
        ("ms_deposit", ctypes.c_double),
        ("ms_fft", ctypes.c_double),
        ("ms_interp", ctypes.c_double),
        ("ms_tree", ctypes.c_double),
        ("ms_walk", ctypes.c_double),
        ("interactions", ctypes.c_uint64),
        ("nodes", ctypes.c_uint64),
        ("groups", ctypes.c_uint64),
        ("listed", ctypes.c_uint64),
        ("kept", ctypes.c_uint64),
        ("ms_walk_kernel", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(ctypes.Structure):
    _fields_ = [("fft_cap", ctypes.c_int32), ("group_level_shift", ctypes.c_int32), ("fft_path", ctypes.c_int32),
                ("tree_depth", ctypes.c_int32), ("small_node_direct", ctypes.c_int32), ("ms_default", ctypes.c_double),
                ("ms_tuned", ctypes.c_double), ("err_default", ctypes.c_double), ("err_tuned", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class EnsembleParams(ctypes.Structure):
    _fields_ = [("box_size", ctypes.c_double), ("omega_m", ctypes.c_double),
                ("softening_scale", ctypes.c_double), ("theta", ctypes.c_double)]


class RunSettings(ctypes.Structure):
    """RunSettings (src/runner.rs:29-47)."""
    _fields_ = [("z_start", ctypes.c_double), ("z_end", ctypes.c_double), ("steps", ctypes.c_int32),
                ("particles_per_side", ctypes.c_int32), ("seed", ctypes.c_uint64), ("treepm", ctypes.c_int32),
                ("theta", ctypes.c_double), ("softening_scale", ctypes.c_double), ("pacing_ms", ctypes.c_uint32)]

    @classmethod
    def default(cls):
        s = cls()
        load().tpm_run_settings_default(ctypes.byref(s))
        return s

    def sanitized(self):
        out = RunSettings()
        rc = load().tpm_run_settings_sanitized(ctypes.byref(self), ctypes.byref(out))
        if rc != TPM_OK:
            raise TpmError(rc, "settings are not finite")
        return out

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SnapshotHeader(ctypes.Structure):
    """tpm_snapshot_header: the 128-byte header of a binary snapshot file."""
    _fields_ = [("magic", ctypes.c_char * 8), ("version", ctypes.c_uint32), ("header_bytes", ctypes.c_uint32),
                ("n", ctypes.c_uint64), ("a", ctypes.c_double), ("box_size", ctypes.c_double),
                ("omega_m", ctypes.c_double), ("omega_l", ctypes.c_double), ("omega_r", ctypes.c_double),
                ("h0", ctypes.c_double), ("softening", ctypes.c_double), ("r_split", ctypes.c_double),
                ("r_cut", ctypes.c_double), ("theta", ctypes.c_double), ("mesh_n", ctypes.c_int32),
                ("step", ctypes.c_int32), ("seed", ctypes.c_uint64), ("reserved", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Frame(ctypes.Structure):
    _fields_ = [("step", ctypes.c_int32), ("steps", ctypes.c_int32), ("redshift", ctypes.c_double),
                ("max_speed", ctypes.c_double), ("elapsed_s", ctypes.c_double), ("n", ctypes.c_size_t),
                ("d_pos", ctypes.c_void_p), ("d_speed", ctypes.c_void_p), ("d_density", ctypes.c_void_p),
                ("d_rsd", ctypes.c_void_p), ("d_delta", ctypes.c_void_p), ("mesh_n", ctypes.c_int32),
                ("a", ctypes.c_double), ("d_state_pos", ctypes.c_void_p), ("d_state_mom", ctypes.c_void_p),
                ("snapshot", ctypes.POINTER(SnapshotHeader))]


FRAME_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(Frame))

_lib = None


def load():
    """Loads libtpm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libtpm.so not built at {LIB_PATH}; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    P, S, D, F = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double, ctypes.c_float
    sigs = {
        "tpm_version": (ctypes.c_char_p, []),
        "tpm_create": (ctypes.c_int, [ctypes.POINTER(Params), S, ctypes.POINTER(ctypes.c_void_p)]),
        "tpm_destroy": (None, [P]),
        "tpm_last_error": (ctypes.c_char_p, [P]),
        "tpm_set_params": (ctypes.c_int, [P, ctypes.POINTER(Params)]),
        "tpm_accelerations_f32": (ctypes.c_int, [P, P, S, D, P, P]),
        "tpm_accelerations_f64": (ctypes.c_int, [P, P, S, D, P, P]),
        "tpm_accelerations_dev_f64": (ctypes.c_int, [P, P, S, D, P, P, P]),
        "tpm_get_stats": (ctypes.c_int, [P, ctypes.POINTER(Stats)]),
        "tpm_set_timing": (ctypes.c_int, [P, ctypes.c_int]),
        "tpm_set_counting": (ctypes.c_int, [P, ctypes.c_int]),
        "tpm_set_overlap": (ctypes.c_int, [P, ctypes.c_int]),
        "tpm_deposit_f32": (ctypes.c_int, [P, P, S, F, P]),
        "tpm_deposit_f64": (ctypes.c_int, [P, P, S, D, P]),
        "tpm_density_contrast_f32": (ctypes.c_int, [P, P, S, P]),
        "tpm_pm_accel": (ctypes.c_int, [P, P, D, D, P, P, P]),
        "tpm_interpolate_f32": (ctypes.c_int, [P, P, P, S, P]),
        "tpm_short_range_f32": (ctypes.c_int, [P, P, S, D, P]),
        "tpm_short_range_f64": (ctypes.c_int, [P, P, S, D, P]),
        "tpm_judge_f64": (ctypes.c_int, [P, P, S, P, S, D, D, P]),
        "tpm_zeldovich": (ctypes.c_int, [P, ctypes.c_uint64, D, D, D, P, P]),
        "tpm_zeldovich_dev": (ctypes.c_int, [P, ctypes.c_uint64, D, D, D, P, P, P]),
        "tpm_kdk_factors": (ctypes.c_int, [ctypes.POINTER(Cosmology), D, D, P, P, P, P]),
        "tpm_step_to_dev": (ctypes.c_int, [P, ctypes.POINTER(Cosmology), P, P, S, D, D, P, P, P]),
        "tpm_run": (ctypes.c_int, [P, ctypes.POINTER(Cosmology), P, P, S, D, D, ctypes.c_int]),
        "tpm_device_alloc": (ctypes.c_int, [S, ctypes.POINTER(ctypes.c_void_p)]),
        "tpm_device_free": (ctypes.c_int, [P]),
        "tpm_memcpy": (ctypes.c_int, [P, P, S, ctypes.c_int]),
        "tpm_sync": (ctypes.c_int, [P]),
        "tpm_nfw_halos": (ctypes.c_int, [P, ctypes.POINTER(NfwSpec), S, P]),
        "tpm_nfw_halos_dev": (ctypes.c_int, [P, ctypes.POINTER(NfwSpec), S, P, P]),
        "tpm_mlp_param_count": (S, []),
        "tpm_mlp_init": (ctypes.c_int, [P, ctypes.c_uint64]),
        "tpm_mlp_get_params": (ctypes.c_int, [P, P, S]),
        "tpm_mlp_set_params": (ctypes.c_int, [P, P, S]),
        "tpm_mlp_gradients": (ctypes.c_int, [P, P, P, S, P, ctypes.POINTER(ctypes.c_double)]),
        "tpm_mlp_predict": (ctypes.c_int, [P, P, S, P]),
        "tpm_mlp_train_dev": (ctypes.c_int, [P, P, S, D, S, ctypes.c_int, S, F, ctypes.c_uint64, P]),
        "tpm_accelerations_learned_dev": (ctypes.c_int, [P, P, S, D, D, P, ctypes.POINTER(ctypes.c_uint64), P]),
        "tpm_mlp_error_scale": (ctypes.c_int, [P, D, ctypes.POINTER(ctypes.c_double)]),
        "tpm_growth": (ctypes.c_int, [ctypes.POINTER(Cosmology), D, P, P, P]),
        "tpm_power_spectrum": (ctypes.c_int, [ctypes.POINTER(Cosmology), D, D, D, D, P, P]),
        "tpm_power_sigma": (ctypes.c_int, [ctypes.POINTER(Cosmology), D, D, D, D, P]),
        "tpm_zeldovich_bbks": (ctypes.c_int, [P, ctypes.POINTER(Cosmology), D, D, D, ctypes.c_uint64, D, P, P]),
        "tpm_zeldovich_bbks_dev": (ctypes.c_int, [P, ctypes.POINTER(Cosmology), D, D, D, ctypes.c_uint64, D, P, P,
                                                  P]),
        "tpm_run_settings_default": (None, [ctypes.POINTER(RunSettings)]),
        "tpm_run_settings_sanitized": (ctypes.c_int, [ctypes.POINTER(RunSettings), ctypes.POINTER(RunSettings)]),
        "tpm_headless_run": (ctypes.c_int, [ctypes.POINTER(RunSettings), FRAME_FN, P, P, P, S]),
        "tpm_snapshot_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(SnapshotHeader), P, P]),
        "tpm_snapshot_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(SnapshotHeader), P, P, S]),
        "tpm_snapshot_dev": (ctypes.c_int, [P, P, P, S, D, D, P, P, P, P, P, P]),
        "tpm_downsample_grid_dev": (ctypes.c_int, [P, P, P, ctypes.POINTER(ctypes.c_int32), P]),
        "tpm_autotune": (ctypes.c_int, [P, P, S, D, P, ctypes.POINTER(Tuning)]),
        "tpm_set_tuning": (ctypes.c_int, [P, ctypes.POINTER(Tuning)]),
        "tpm_ensemble_create": (ctypes.c_int, [ctypes.POINTER(EnsembleParams), ctypes.c_int, P,
                                               ctypes.POINTER(ctypes.c_void_p)]),
        "tpm_ensemble_destroy": (None, [P]),
        "tpm_ensemble_last_error": (ctypes.c_char_p, [P]),
        "tpm_ensemble_ics": (ctypes.c_int, [P, P, D, D, D]),
        "tpm_ensemble_autotune": (ctypes.c_int, [P, D, P]),
        "tpm_ensemble_run": (ctypes.c_int, [P, ctypes.POINTER(Cosmology), D, D, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_double)]),
        "tpm_ensemble_get": (ctypes.c_int, [P, ctypes.c_int, P, P, P]),
        "tpm_deposit_dev_f64_noalloc": (ctypes.c_int, [P, S, ctypes.c_int32, D, D, P, P]),
        "tpm_get_device_status": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int]),
        "tpm_morton_order": (ctypes.c_int, [P, P, S, P, P]),
        "tpm_ensemble_graph_nodes": (ctypes.c_int, [P, P, P, P]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(shape) if shape is not None else a


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


def force_split(box_size, mesh_n, particles_per_side, softening_scale, theta):
    """ForceSplit::with (src/tree.rs:40-50): r_s = 1.25 cells, r_cut = 4.5 r_s, FP32 like the reference."""
    f = np.float32
    cell = f(box_size) / f(mesh_n)
    r_split = f(1.25) * cell
    spacing = f(box_size) / f(particles_per_side)
    return {"r_split": float(r_split), "r_cut": float(f(4.5) * r_split),
            "softening": float(f(softening_scale) * spacing), "theta": float(theta)}


class Solver:
    """One libtpm context: mesh, FFT tables, tree scratch and device buffers."""

    def __init__(self, mesh_n, box_size, omega_m=0.315, split=None, max_particles=1 << 16, flags=0):
        self.lib = load()
        p = Params()
        p.mesh_n = int(mesh_n)
        p.box_size = float(box_size)
        p.omega_m = float(omega_m)
        fl = int(flags)
        if split is not None:
            fl |= FLAG_TREEPM
            p.r_split = float(split["r_split"])
            p.r_cut = float(split["r_cut"])
            p.softening = float(split["softening"])
            p.theta = float(split["theta"])
        p.flags = fl
        self.params = p
        self.n = int(mesh_n)
        self.max_particles = int(max_particles)
        h = ctypes.c_void_p()
        rc = self.lib.tpm_create(ctypes.byref(p), self.max_particles, ctypes.byref(h))
        if rc != TPM_OK:
            raise TpmError(rc, self.lib.tpm_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.tpm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != TPM_OK:
            raise TpmError(rc, self.lib.tpm_last_error(self.h).decode())

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        self._check(self.lib.tpm_set_params(self.h, ctypes.byref(self.params)))

    def set_timing(self, on=True):
        self._check(self.lib.tpm_set_timing(self.h, 1 if on else 0))

    def set_counting(self, on=True):
        """Interaction counting in the walk (on by default); off, stats()["interactions"] reads 0."""
        self._check(self.lib.tpm_set_counting(self.h, 1 if on else 0))

    def set_overlap(self, on=True):
        """The octree on a second stream beside the PM part above 2^22 particles (on by default)."""
        self._check(self.lib.tpm_set_overlap(self.h, 1 if on else 0))

    def stats(self):
        s = Stats()
        self._check(self.lib.tpm_get_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def accelerations(self, pos, a, want_delta=False):
        """Simulation::accelerations: returns (acc (N,3) float32, delta or None)."""
        if np.asarray(pos).dtype == np.float64:
            pos = _f64(pos, (-1, 3))
            fn = self.lib.tpm_accelerations_f64
        else:
            pos = _f32(pos, (-1, 3))
            fn = self.lib.tpm_accelerations_f32
        acc = np.empty((len(pos), 3), np.float32)
        delta = np.empty(self.n ** 3, np.float32) if want_delta else None
        self._check(fn(self.h, _ptr(pos), len(pos), float(a), _ptr(acc), _ptr(delta)))
        return acc, delta

    def accelerations_dev(self, d_pos, n, a, d_acc, d_delta=0, stream=0):
        """Device-pointer form (integers are device addresses, e.g. torch's data_ptr())."""
        self._check(self.lib.tpm_accelerations_dev_f64(self.h, ctypes.c_void_p(d_pos), int(n), float(a),
                                                       ctypes.c_void_p(d_acc), ctypes.c_void_p(d_delta or None),
                                                       ctypes.c_void_p(stream or None)))

    def deposit(self, pos, mass=1.0):
        """Mesh::deposit of `mass` per particle; returns the n^3 grid (float32)."""
        out = np.empty(self.n ** 3, np.float32)
        if np.asarray(pos).dtype == np.float64:
            pos = _f64(pos, (-1, 3))
            self._check(self.lib.tpm_deposit_f64(self.h, _ptr(pos), len(pos), float(mass), _ptr(out)))
        else:
            pos = _f32(pos, (-1, 3))
            self._check(self.lib.tpm_deposit_f32(self.h, _ptr(pos), len(pos), float(mass), _ptr(out)))
        return out

    def density_contrast(self, pos):
        pos = _f32(pos, (-1, 3))
        out = np.empty(self.n ** 3, np.float32)
        self._check(self.lib.tpm_density_contrast_f32(self.h, _ptr(pos), len(pos), _ptr(out)))
        return out

    def pm_accel(self, delta, a, r_split=0.0):
        """PoissonSolver::acceleration: returns (3, n^3) float32 grids."""
        delta = _f32(delta).reshape(-1)
        assert delta.size == self.n ** 3, "mesh size must match the solver"
        g = np.empty((3, self.n ** 3), np.float32)
        self._check(self.lib.tpm_pm_accel(self.h, _ptr(delta), float(a), float(r_split),
                                          _ptr(g[0]), _ptr(g[1]), _ptr(g[2])))
        return g

    def interpolate(self, cells, pos):
        cells = _f32(cells).reshape(-1)
        pos = _f32(pos, (-1, 3))
        out = np.empty(len(pos), np.float32)
        self._check(self.lib.tpm_interpolate_f32(self.h, _ptr(cells), _ptr(pos), len(pos), _ptr(out)))
        return out

    def short_range(self, pos, k):
        """tree::accelerations (short-range force only) with coupling k."""
        if np.asarray(pos).dtype == np.float64:
            pos = _f64(pos, (-1, 3))
            fn = self.lib.tpm_short_range_f64
        else:
            pos = _f32(pos, (-1, 3))
            fn = self.lib.tpm_short_range_f32
        acc = np.empty((len(pos), 3), np.float32)
        self._check(fn(self.h, _ptr(pos), len(pos), float(k), _ptr(acc)))
        return acc

    def zeldovich(self, seed, k_c, disp_rms, mom_factor=0.0, want_momenta=False):
        """Lattice Zel'dovich ICs (mesh_n^3 particles): returns (pos (N,3) f64, mom or None)."""
        npart = self.n ** 3
        pos = np.empty((npart, 3), np.float64)
        mom = np.empty((npart, 3), np.float64) if want_momenta else None
        self._check(self.lib.tpm_zeldovich(self.h, int(seed), float(k_c), float(disp_rms), float(mom_factor),
                                           _ptr(pos), _ptr(mom)))
        return pos, mom

    def zeldovich_bbks(self, seed, a_init, cosmo=None, h0=H0, n_s=N_S, sigma8=SIGMA8):
        """zeldovich (src/ics.rs:111-182) with the sigma_8-normalised BBKS spectrum: (pos, mom), FP64."""
        cosmo = cosmo or Cosmology.lcdm_default()
        npart = self.n ** 3
        pos = np.empty((npart, 3), np.float64)
        mom = np.empty((npart, 3), np.float64)
        self._check(self.lib.tpm_zeldovich_bbks(self.h, ctypes.byref(cosmo), float(h0), float(n_s), float(sigma8),
                                                int(seed), float(a_init), _ptr(pos), _ptr(mom)))
        return pos, mom

    def zeldovich_bbks_dev(self, seed, a_init, d_pos, d_mom=0, stream=0, cosmo=None, h0=H0, n_s=N_S, sigma8=SIGMA8):
        cosmo = cosmo or Cosmology.lcdm_default()
        self._check(self.lib.tpm_zeldovich_bbks_dev(self.h, ctypes.byref(cosmo), float(h0), float(n_s),
                                                    float(sigma8), int(seed), float(a_init), ctypes.c_void_p(d_pos),
                                                    ctypes.c_void_p(d_mom or None), ctypes.c_void_p(stream or None)))

    def zeldovich_dev(self, seed, k_c, disp_rms, mom_factor, d_pos, d_mom=0, stream=0):
        self._check(self.lib.tpm_zeldovich_dev(self.h, int(seed), float(k_c), float(disp_rms), float(mom_factor),
                                               ctypes.c_void_p(d_pos), ctypes.c_void_p(d_mom or None),
                                               ctypes.c_void_p(stream or None)))

    def run(self, pos, mom, a_start, a_end, steps, cosmo=None):
        """Simulation::new + run: returns (pos, mom) after `steps` KDK steps (FP64 on the GPU)."""
        cosmo = cosmo or Cosmology.lcdm_default()
        pos = np.array(pos, dtype=np.float64, copy=True).reshape(-1, 3)
        mom = np.array(mom, dtype=np.float64, copy=True).reshape(-1, 3)
        self._check(self.lib.tpm_run(self.h, ctypes.byref(cosmo), _ptr(pos), _ptr(mom), len(pos), float(a_start),
                                     float(a_end), int(steps)))
        return pos, mom

    def step_to_dev(self, d_pos, d_mom, n, a0, a1, d_acc, d_delta=0, stream=0, cosmo=None):
        cosmo = cosmo or Cosmology.lcdm_default()
        self._check(self.lib.tpm_step_to_dev(self.h, ctypes.byref(cosmo), ctypes.c_void_p(d_pos),
                                             ctypes.c_void_p(d_mom), int(n), float(a0), float(a1),
                                             ctypes.c_void_p(d_acc), ctypes.c_void_p(d_delta or None),
                                             ctypes.c_void_p(stream or None)))

    def nfw_halos(self, spec, n):
        pos = np.empty((int(n), 3), np.float64)
        self._check(self.lib.tpm_nfw_halos(self.h, ctypes.byref(spec), int(n), _ptr(pos)))
        return pos

    def nfw_halos_dev(self, spec, n, d_pos, stream=0):
        self._check(self.lib.tpm_nfw_halos_dev(self.h, ctypes.byref(spec), int(n), ctypes.c_void_p(d_pos),
                                               ctypes.c_void_p(stream or None)))

    # ------------------------------------------------------------ learned correction
    def mlp_init(self, seed=1):
        self._check(self.lib.tpm_mlp_init(self.h, int(seed)))

    def mlp_params(self):
        n = self.lib.tpm_mlp_param_count()
        out = np.empty(n, np.float32)
        self._check(self.lib.tpm_mlp_get_params(self.h, _ptr(out), n))
        return out

    def mlp_set_params(self, params):
        params = _f32(params).reshape(-1)
        self._check(self.lib.tpm_mlp_set_params(self.h, _ptr(params), params.size))

    def mlp_gradients(self, x, t):
        x = _f32(x, (-1, 33))
        t = _f32(t, (-1, 3))
        g = np.empty(self.lib.tpm_mlp_param_count(), np.float32)
        loss = ctypes.c_double()
        self._check(self.lib.tpm_mlp_gradients(self.h, _ptr(x), _ptr(t), len(x), _ptr(g), ctypes.byref(loss)))
        return g, loss.value

    def mlp_predict(self, x):
        x = _f32(x, (-1, 33))
        out = np.empty((len(x), 4), np.float32)
        self._check(self.lib.tpm_mlp_predict(self.h, _ptr(x), len(x), _ptr(out)))
        return out

    def mlp_train_dev(self, d_pos, n, a, n_samples, epochs, batch=8192, lr=1e-3, seed=1):
        hist = np.zeros(max(int(epochs), 1), np.float64)
        self._check(self.lib.tpm_mlp_train_dev(self.h, ctypes.c_void_p(d_pos), int(n), float(a), int(n_samples),
                                               int(epochs), int(batch), float(lr), int(seed), _ptr(hist)))
        return hist[:int(epochs)]

    def accelerations_learned_dev(self, d_pos, n, a, tol, d_acc, d_walked=0):
        """Learned TreePM force; returns the number of walked particles (d_walked: optional device mask)."""
        nf = ctypes.c_uint64()
        self._check(self.lib.tpm_accelerations_learned_dev(self.h, ctypes.c_void_p(d_pos), int(n), float(a),
                                                           float(tol), ctypes.c_void_p(d_acc), ctypes.byref(nf),
                                                           ctypes.c_void_p(d_walked or None)))
        return int(nf.value)

    def mlp_error_scale(self, tol):
        """The gate's calibrated error-estimate multiplier kappa for tolerance tol (inf: accepts nothing)."""
        v = ctypes.c_double()
        self._check(self.lib.tpm_mlp_error_scale(self.h, float(tol), ctypes.byref(v)))
        return v.value

    def sync(self):
        self._check(self.lib.tpm_sync(self.h))

    def morton_order(self, pos):
        """Octree (Morton) order of FP64 positions: (order (N,) uint32, sorted 63-bit keys (N,) uint64)."""
        pos = _f64(pos, (-1, 3))
        order = np.empty(len(pos), np.uint32)
        keys = np.empty(len(pos), np.uint64)
        self._check(self.lib.tpm_morton_order(self.h, _ptr(pos), len(pos), _ptr(order), _ptr(keys)))
        return order, keys

    def device_status(self, clear=False):
        """Sticky device-side failure bits (STATUS_*) of the evaluations since the last clear."""
        v = ctypes.c_uint32()
        self._check(self.lib.tpm_get_device_status(self.h, ctypes.byref(v), 1 if clear else 0))
        return int(v.value)

    def snapshot_dev(self, d_pos, d_mom, n, a, e, d_delta, d_pos32=0, d_speed=0, d_density=0, d_rsd=0, stream=0):
        """snapshot (src/runner.rs:262-284) into device buffers (0 = not wanted)."""
        v = lambda x: ctypes.c_void_p(x or None)  # noqa: E731
        self._check(self.lib.tpm_snapshot_dev(self.h, v(d_pos), v(d_mom), int(n), float(a), float(e), v(d_delta),
                                              v(d_pos32), v(d_speed), v(d_density), v(d_rsd), v(stream)))

    def downsample_grid_dev(self, d_delta, d_out=0, stream=0):
        """downsample_grid (src/frame.rs:33-52); returns the output side m."""
        m = ctypes.c_int32()
        self._check(self.lib.tpm_downsample_grid_dev(self.h, ctypes.c_void_p(d_delta or None),
                                                     ctypes.c_void_p(d_out or None), ctypes.byref(m),
                                                     ctypes.c_void_p(stream or None)))
        return m.value

    def autotune(self, d_pos, n, a, d_acc):
        """Times the FFT and walk-group knobs on this input and keeps the fastest (returns the choice)."""
        t = Tuning()
        self._check(self.lib.tpm_autotune(self.h, ctypes.c_void_p(d_pos), int(n), float(a), ctypes.c_void_p(d_acc),
                                          ctypes.byref(t)))
        return t

    def set_tuning(self, fft_cap=16, group_level_shift=0, fft_path=0, tree_depth=21, small_node_direct=0):
        t = Tuning()
        t.fft_cap, t.group_level_shift, t.fft_path = int(fft_cap), int(group_level_shift), int(fft_path)
        t.tree_depth, t.small_node_direct = int(tree_depth), int(small_node_direct)
        self._check(self.lib.tpm_set_tuning(self.h, ctypes.byref(t)))

    def judge(self, pos, idx, a, theta=0.2):
        """FP64 TreePM force (FP64 PM + FP64 per-particle walk at `theta`) at caller indices idx."""
        pos = _f64(pos, (-1, 3))
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        out = np.empty((len(idx), 3), np.float64)
        self._check(self.lib.tpm_judge_f64(self.h, _ptr(pos), len(pos), _ptr(idx), len(idx), float(a),
                                           float(theta), _ptr(out)))
        return out


class Ensemble:
    """Independent periodic boxes (box b: sides[b]^3 particles on a sides[b]^3 mesh), each run
    from its own Zel'dovich ICs through KDK steps; a run can be one HIP graph launch."""

    def __init__(self, sides, box_size=100.0, omega_m=None, softening_scale=1 / 30, theta=0.5):
        self.lib = load()
        self.sides = [int(s) for s in sides]
        p = EnsembleParams()
        p.box_size = float(box_size)
        p.omega_m = float(Cosmology.lcdm_default().omega_m if omega_m is None else omega_m)
        p.softening_scale = float(softening_scale)
        p.theta = float(theta)
        arr = np.ascontiguousarray(self.sides, dtype=np.int32)
        h = ctypes.c_void_p()
        rc = self.lib.tpm_ensemble_create(ctypes.byref(p), len(arr), _ptr(arr), ctypes.byref(h))
        if rc != TPM_OK:
            raise TpmError(rc, self.lib.tpm_last_error(None).decode())
        self.h = h

    def _check(self, rc):
        if rc != TPM_OK:
            raise TpmError(rc, self.lib.tpm_ensemble_last_error(self.h).decode())

    def ics(self, seeds, k_c=0.25, disp_cells=0.3, mom_factor=0.0):
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        assert len(seeds) == len(self.sides)
        self._check(self.lib.tpm_ensemble_ics(self.h, _ptr(seeds), float(k_c), float(disp_cells), float(mom_factor)))

    def autotune(self, a):
        out = (Tuning * len(self.sides))()
        self._check(self.lib.tpm_ensemble_autotune(self.h, float(a), out))
        return [t.as_dict() for t in out]

    def run(self, a_start, a_end, steps, graph=True, cosmo=None):
        """Initial force plus `steps` KDK steps of every box; returns the run's device milliseconds."""
        cosmo = cosmo or Cosmology.lcdm_default()
        ms = ctypes.c_double()
        self._check(self.lib.tpm_ensemble_run(self.h, ctypes.byref(cosmo), float(a_start), float(a_end), int(steps),
                                              1 if graph else 0, ctypes.byref(ms)))
        return ms.value

    def graph_nodes(self):
        """Node counts of the captured run's graph: {"kernels", "memsets", "others"}."""
        v = [ctypes.c_int32() for _ in range(3)]
        self._check(self.lib.tpm_ensemble_graph_nodes(self.h, *[ctypes.byref(x) for x in v]))
        return {"kernels": v[0].value, "memsets": v[1].value, "others": v[2].value}

    def get(self, b):
        n = self.sides[b] ** 3
        pos = np.empty((n, 3), np.float64)
        mom = np.empty((n, 3), np.float64)
        acc = np.empty((n, 3), np.float32)
        self._check(self.lib.tpm_ensemble_get(self.h, int(b), _ptr(pos), _ptr(mom), _ptr(acc)))
        return pos, mom, acc

    def close(self):
        if self.h:
            self.lib.tpm_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def headless_run(settings, on_frame=None, want_state=False):
    """The `--no-viz` run (src/runner.rs:186-260): calls on_frame(frame) after every step (frame
    fields as in tpm_frame; device pointers valid during the call; return True to stop). Returns
    (pos, mom) of the final state when want_state, else None."""
    lib = load()
    s = settings.sanitized()
    npart = s.particles_per_side ** 3
    errors = []

    def cb(_user, fp):
        try:
            return 1 if (on_frame is not None and on_frame(fp.contents)) else 0
        except Exception as e:  # surfaced after the run
            errors.append(e)
            return 1

    fn = FRAME_FN(cb)
    pos = np.empty((npart, 3), np.float64) if want_state else None
    mom = np.empty((npart, 3), np.float64) if want_state else None
    rc = lib.tpm_headless_run(ctypes.byref(settings), fn, None, _ptr(pos), _ptr(mom), npart if want_state else 0)
    if errors:
        raise errors[0]
    if rc != TPM_OK:
        raise TpmError(rc, lib.tpm_last_error(None).decode())
    return (pos, mom) if want_state else None


def snapshot_write(path, header, pos, mom):
    """Writes a binary snapshot (tpm_snapshot_write): header (SnapshotHeader; n is taken from pos),
    positions and momenta (n x 3 float64, caller order)."""
    pos, mom = _f64(pos, (-1, 3)), _f64(mom, (-1, 3))
    if pos.shape != mom.shape:
        raise ValueError("positions and momenta differ in shape")
    h = SnapshotHeader.from_buffer_copy(header)
    h.n = len(pos)
    rc = load().tpm_snapshot_write(os.fsencode(path), ctypes.byref(h), _ptr(pos), _ptr(mom))
    if rc != TPM_OK:
        raise TpmError(rc, f"cannot write snapshot {path}")


def snapshot_read(path, arrays=True):
    """Reads a binary snapshot: (header, pos, mom), or the header alone when arrays is False."""
    lib = load()
    h = SnapshotHeader()
    rc = lib.tpm_snapshot_read(os.fsencode(path), ctypes.byref(h), None, None, 0)
    if rc != TPM_OK:
        raise TpmError(rc, f"cannot read snapshot {path}")
    if not arrays:
        return h
    pos = np.empty((h.n, 3), np.float64)
    mom = np.empty((h.n, 3), np.float64)
    rc = lib.tpm_snapshot_read(os.fsencode(path), ctypes.byref(h), _ptr(pos), _ptr(mom), int(h.n))
    if rc != TPM_OK:
        raise TpmError(rc, f"cannot read snapshot {path}")
    return h, pos, mom


def coupling(omega_m, box_size, a, n):
    """K = 1.5 Omega_m L^3 / (4 pi a N) in float32 (src/integrator.rs:150-152)."""
    f = np.float32
    l = f(box_size)
    return float(f(1.5) * f(omega_m) * (l * l * l) / (f(4.0) * f(np.pi) * f(a) * f(n)))


def kdk_factors(a0, a1, cosmo=None):
    """(am, k1, d, k2) of one KDK step (src/integrator.rs:174-186, 262-297), FP64."""
    cosmo = cosmo or Cosmology.lcdm_default()
    out = [ctypes.c_double() for _ in range(4)]
    rc = load().tpm_kdk_factors(ctypes.byref(cosmo), float(a0), float(a1), *[ctypes.byref(o) for o in out])
    if rc != TPM_OK:
        raise TpmError(rc, "invalid KDK factor arguments")
    return tuple(o.value for o in out)


def growth(a, cosmo=None):
    """(D(a), f(a), E(a)): Cosmology::growth, growth_rate and e (src/cosmology.rs:31-91), FP64."""
    cosmo = cosmo or Cosmology.lcdm_default()
    out = [ctypes.c_double() for _ in range(3)]
    rc = load().tpm_growth(ctypes.byref(cosmo), float(a), *[ctypes.byref(o) for o in out])
    if rc != TPM_OK:
        raise TpmError(rc, "invalid growth arguments")
    return tuple(o.value for o in out)


def power_spectrum(k, cosmo=None, h0=H0, n_s=N_S, sigma8=SIGMA8):
    """(P(k), amplitude) of the sigma_8-normalised BBKS spectrum (src/ics.rs:20-64), FP64."""
    cosmo = cosmo or Cosmology.lcdm_default()
    p, amp = ctypes.c_double(), ctypes.c_double()
    rc = load().tpm_power_spectrum(ctypes.byref(cosmo), float(h0), float(n_s), float(sigma8), float(k),
                                   ctypes.byref(p), ctypes.byref(amp))
    if rc != TPM_OK:
        raise TpmError(rc, "invalid spectrum arguments")
    return p.value, amp.value


def power_sigma(radius, cosmo=None, h0=H0, n_s=N_S, sigma8=SIGMA8):
    """sigma(R) of the normalised spectrum (PowerSpectrum::sigma, src/ics.rs:35-52)."""
    cosmo = cosmo or Cosmology.lcdm_default()
    out = ctypes.c_double()
    rc = load().tpm_power_sigma(ctypes.byref(cosmo), float(h0), float(n_s), float(sigma8), float(radius),
                                ctypes.byref(out))
    if rc != TPM_OK:
        raise TpmError(rc, "invalid spectrum arguments")
    return out.value


def zeldovich_momentum_factor(a, cosmo=None):
    """p = a^2 f(a) E(a) x displacement at a (src/ics.rs:155-160, the displacement already at a)."""
    _, f, e = growth(a, cosmo)
    return a * a * f * e


class DeviceArray:
    """A device allocation made by libtpm's own HIP runtime (for the *_dev entry points)."""

    def __init__(self, count, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = ctypes.c_void_p()
        rc = load().tpm_device_alloc(self.nbytes, ctypes.byref(p))
        if rc != TPM_OK:
            raise TpmError(rc, f"device allocation of {self.nbytes} bytes failed")
        self.ptr = p.value or 0

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.count
        rc = load().tpm_memcpy(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p), self.nbytes, 0)
        if rc != TPM_OK:
            raise TpmError(rc, "host to device copy failed")

    def numpy(self):
        out = np.empty(self.count, self.dtype)
        rc = load().tpm_memcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr), self.nbytes, 1)
        if rc != TPM_OK:
            raise TpmError(rc, "device to host copy failed")
        return out

    def free(self):
        if self.ptr:
            load().tpm_device_free(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
