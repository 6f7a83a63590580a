"""ctypes loader for the CPU oracle (liboracle.so, built from treepm_oracle.c).

The oracle is test infrastructure: it restates the reference TreePM path
(src/mesh.rs, src/fft.rs, src/poisson.rs, src/tree.rs, src/integrator.rs:117-161 of the
reference crate) on the CPU. Only tests, the smoke check and the benchmark's CPU-baseline
leg import this module, and only as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int


def build():
    """Compile liboracle.so with the committed Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        for sfx, R in (("_f32", ctypes.c_float), ("_f64", ctypes.c_double)):
            def sig(name, res, args):
                f = getattr(L, name + sfx)
                f.restype = res
                f.argtypes = args
            sig("ora_deposit", None, [_P, _I64, _I, R, R, _P])
            sig("ora_density_contrast", None, [_P, _I64, _I, R, _P])
            sig("ora_interpolate", None, [_P, _I, R, _P, _I64, _P])
            sig("ora_wavenumber", R, [_I64, _I64, R])
            sig("ora_fft1d", None, [_P, _I, _I])
            sig("ora_fft3", None, [_P, _I, _I])
            sig("ora_poisson_accel", None, [_P, _I, R, R, R, R, _P, _P, _P])
            sig("ora_short_range_factor", R, [R, R])
            sig("ora_tree_accel", _I64, [_P, _I64, R, R, R, R, R, R, _P])
            sig("ora_tree_accel_subset", None, [_P, _I64, _P, _I64, R, R, R, R, R, R, _P])
            sig("ora_brute_force", None, [_P, _I64, R, R, R, R, R, _P])
            sig("ora_coupling", R, [R, R, R, _I64])
            sig("ora_accelerations", _I64, [_P, _I64, _I, R, R, R, _I, R, R, R, R, _P, _P])
            sig("ora_cosmo_e", R, [R, R, R, R])
            sig("ora_growth", R, [R, R, R, R])
            sig("ora_growth_rate", R, [R, R, R, R])
            sig("ora_kick_factor", R, [R, R, R, R, R])
            sig("ora_drift_factor", R, [R, R, R, R, R])
            sig("ora_midpoint", R, [R, R, R, R, R])
            sig("ora_sim_run", None, [_P, _P, _I64, _I, R, R, R, R, R, R, _I, _I, R, R, R, R])
        L.ora_threads.restype = _I
        L.ora_set_threads.argtypes = [_I]
        L.ora_downsample_grid.restype = _I
        L.ora_downsample_grid.argtypes = [_P, _I, _P]
        L.ora_snapshot_f64.restype = None
        L.ora_snapshot_f64.argtypes = [_P, _P, _I64, ctypes.c_double, ctypes.c_double, _P, _I, ctypes.c_double,
                                       _P, _P, _P, _P]
        _lib = L
    return _lib


def _dt(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "_f32", np.float32
    if dtype == np.float64:
        return "_f64", np.float64
    raise TypeError(dtype)


def _arr(x, dt):
    return np.ascontiguousarray(x, dtype=dt)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def threads():
    return lib().ora_threads()


def set_threads(t):
    lib().ora_set_threads(int(t))


def deposit(pos, n, box_size, mass=1.0, dtype=np.float32):
    """Mesh::deposit (mesh.rs:62-69); returns the n^3 grid, index (ix*n+iy)*n+iz."""
    s, dt = _dt(dtype)
    pos = _arr(pos, dt).reshape(-1, 3)
    out = np.empty(n * n * n, dt)
    getattr(lib(), "ora_deposit" + s)(_p(pos), len(pos), n, box_size, mass, _p(out))
    return out


def density_contrast(pos, n, box_size, dtype=np.float32):
    """Simulation::density_contrast (integrator.rs:117-126)."""
    s, dt = _dt(dtype)
    pos = _arr(pos, dt).reshape(-1, 3)
    out = np.empty(n * n * n, dt)
    getattr(lib(), "ora_density_contrast" + s)(_p(pos), len(pos), n, box_size, _p(out))
    return out


def interpolate(cells, n, box_size, pos, dtype=np.float32):
    """Mesh::interpolate (mesh.rs:72-77) at every position."""
    s, dt = _dt(dtype)
    cells = _arr(cells, dt)
    pos = _arr(pos, dt).reshape(-1, 3)
    out = np.empty(len(pos), dt)
    getattr(lib(), "ora_interpolate" + s)(_p(cells), n, box_size, _p(pos), len(pos), _p(out))
    return out


def wavenumber(i, n, box_size, dtype=np.float32):
    s, _ = _dt(dtype)
    return getattr(lib(), "ora_wavenumber" + s)(i, n, box_size)


def fft1d(line, sign, dtype=np.float64):
    s, dt = _dt(dtype)
    c = np.array(line, dtype=np.complex64 if dt == np.float32 else np.complex128, copy=True)
    getattr(lib(), "ora_fft1d" + s)(_p(c), len(c), sign)
    return c


def fft3(grid, sign, dtype=np.float32):
    """Fft3::forward (sign -1) / inverse (sign +1), unnormalised (fft.rs:27-72)."""
    s, dt = _dt(dtype)
    n = grid.shape[0]
    c = np.array(grid, dtype=np.complex64 if dt == np.float32 else np.complex128, copy=True, order="C")
    getattr(lib(), "ora_fft3" + s)(_p(c), n, sign)
    return c


def poisson_accel(delta, n, box_size, omega_m, a, r_split, dtype=np.float32):
    """PoissonSolver::acceleration (poisson.rs:33-88); returns (gx, gy, gz) flat n^3 grids."""
    s, dt = _dt(dtype)
    delta = _arr(delta, dt)
    g = np.empty((3, n * n * n), dt)
    getattr(lib(), "ora_poisson_accel" + s)(_p(delta), n, box_size, omega_m, a, r_split,
                                            _p(g[0]), _p(g[1]), _p(g[2]))
    return g


def short_range_factor(r, r_split, dtype=np.float32):
    s, _ = _dt(dtype)
    return getattr(lib(), "ora_short_range_factor" + s)(r, r_split)


def tree_accel(pos, box_size, r_split, r_cut, softening, theta, k, dtype=np.float32):
    """tree::accelerations (tree.rs:314-323); returns (acc (N,3), interaction count)."""
    s, dt = _dt(dtype)
    pos = _arr(pos, dt).reshape(-1, 3)
    acc = np.empty_like(pos)
    cnt = getattr(lib(), "ora_tree_accel" + s)(_p(pos), len(pos), box_size, r_split, r_cut,
                                               softening, theta, k, _p(acc))
    return acc, cnt


def tree_accel_subset(pos, targets, box_size, r_split, r_cut, softening, theta, k, dtype=np.float32):
    s, dt = _dt(dtype)
    pos = _arr(pos, dt).reshape(-1, 3)
    targets = np.ascontiguousarray(targets, dtype=np.uint32)
    acc = np.empty((len(targets), 3), dt)
    getattr(lib(), "ora_tree_accel_subset" + s)(_p(pos), len(pos), _p(targets), len(targets), box_size,
                                                r_split, r_cut, softening, theta, k, _p(acc))
    return acc


def brute_force(pos, box_size, r_split, r_cut, softening, k, dtype=np.float32):
    """The O(N^2) short-range sum of the reference test (tree.rs:351-370)."""
    s, dt = _dt(dtype)
    pos = _arr(pos, dt).reshape(-1, 3)
    acc = np.empty_like(pos)
    getattr(lib(), "ora_brute_force" + s)(_p(pos), len(pos), box_size, r_split, r_cut, softening, k, _p(acc))
    return acc


def coupling(omega_m, box_size, a, npart, dtype=np.float32):
    """K = 1.5 Omega_m L^3 / (4 pi a N) (integrator.rs:150-152)."""
    s, _ = _dt(dtype)
    return getattr(lib(), "ora_coupling" + s)(omega_m, box_size, a, npart)


def accelerations(pos, n, box_size, omega_m, a, split=None, dtype=np.float32, want_delta=False):
    """Simulation::accelerations (integrator.rs:140-161).

    split is None (pure PM) or a dict with r_split, r_cut, softening, theta.
    Returns (acc (N,3), delta or None, interaction count).
    """
    s, dt = _dt(dtype)
    pos = _arr(pos, dt).reshape(-1, 3)
    acc = np.empty_like(pos)
    delta = np.empty(n * n * n, dt) if want_delta else None
    sp = split or {}
    cnt = getattr(lib(), "ora_accelerations" + s)(
        _p(pos), len(pos), n, box_size, omega_m, a, 1 if split else 0,
        sp.get("r_split", 0.0), sp.get("r_cut", 0.0), sp.get("softening", 0.0), sp.get("theta", 0.0),
        _p(acc), _p(delta) if delta is not None else None)
    return acc, delta, cnt


def force_split(box_size, mesh_n, particles_per_side, softening_scale, theta):
    """ForceSplit::with (tree.rs:40-50), evaluated in float32 like the reference."""
    f = np.float32
    cell = f(box_size) / f(mesh_n)
    r_split = f(1.25) * cell
    spacing = f(box_size) / f(particles_per_side)
    return {"r_split": float(r_split), "r_cut": float(f(4.5) * r_split),
            "softening": float(f(softening_scale) * spacing), "theta": float(theta)}


# Cosmology::lcdm_default (src/cosmology.rs:21-30), evaluated in float32 like the reference.
LCDM = {"omega_r": float(np.float32(9.14e-5)), "omega_m": float(np.float32(0.315)),
        "omega_l": float(np.float32(1.0) - np.float32(0.315) - np.float32(9.14e-5))}


def _cosmo(c):
    c = c or LCDM
    return c["omega_r"], c["omega_m"], c["omega_l"]


def cosmo_e(a, cosmo=None, dtype=np.float32):
    s, _ = _dt(dtype)
    return getattr(lib(), "ora_cosmo_e" + s)(*_cosmo(cosmo), a)


def growth(a, cosmo=None, dtype=np.float32):
    """Cosmology::growth (src/cosmology.rs:82-85)."""
    s, _ = _dt(dtype)
    return getattr(lib(), "ora_growth" + s)(*_cosmo(cosmo), a)


def growth_rate(a, cosmo=None, dtype=np.float32):
    s, _ = _dt(dtype)
    return getattr(lib(), "ora_growth_rate" + s)(*_cosmo(cosmo), a)


def kick_factor(a0, a1, cosmo=None, dtype=np.float32):
    s, _ = _dt(dtype)
    return getattr(lib(), "ora_kick_factor" + s)(*_cosmo(cosmo), a0, a1)


def drift_factor(a0, a1, cosmo=None, dtype=np.float32):
    s, _ = _dt(dtype)
    return getattr(lib(), "ora_drift_factor" + s)(*_cosmo(cosmo), a0, a1)


def midpoint(a0, a1, cosmo=None, dtype=np.float32):
    s, _ = _dt(dtype)
    return getattr(lib(), "ora_midpoint" + s)(*_cosmo(cosmo), a0, a1)


def sim_run(pos, mom, n, box_size, a_start, a_end, steps, split=None, cosmo=None, dtype=np.float32):
    """Simulation::new + run (src/integrator.rs:54-79, 165-186); returns (pos, mom)."""
    s, dt = _dt(dtype)
    pos = np.array(pos, dtype=dt, copy=True).reshape(-1, 3)
    mom = np.array(mom, dtype=dt, copy=True).reshape(-1, 3)
    sp = split or {}
    getattr(lib(), "ora_sim_run" + s)(_p(pos), _p(mom), len(pos), n, box_size, *_cosmo(cosmo), a_start, a_end,
                                      steps, 1 if split else 0, sp.get("r_split", 0.0), sp.get("r_cut", 0.0),
                                      sp.get("softening", 0.0), sp.get("theta", 0.0))
    return pos, mom


def downsample_grid(delta, n):
    """frame.rs:33-52 downsample_grid: (m, grid of m^3 floats, x fastest)."""
    d = _arr(delta, np.float32).reshape(-1)
    m = lib().ora_downsample_grid(_p(d), int(n), None)
    out = np.empty(m ** 3, np.float32)
    lib().ora_downsample_grid(_p(d), int(n), _p(out))
    return m, out


def snapshot(pos, mom, a, e, delta, n, box_size):
    """runner.rs:262-284 snapshot of FP64 state: (positions f32, speeds, densities, rsd)."""
    pos = _arr(pos, np.float64).reshape(-1, 3)
    mom = _arr(mom, np.float64).reshape(-1, 3)
    d = _arr(delta, np.float32).reshape(-1)
    np_ = len(pos)
    opos = np.empty((np_, 3), np.float32)
    ospeed = np.empty(np_, np.float32)
    odens = np.empty(np_, np.float32)
    orsd = np.empty((np_, 3), np.float32)
    lib().ora_snapshot_f64(_p(pos), _p(mom), np_, float(a), float(e), _p(d), int(n), float(box_size), _p(opos),
                           _p(ospeed), _p(odens), _p(orsd))
    return opos, ospeed, odens, orsd
