"""Deterministic inputs of the reference's own tests, restated exactly, plus error metrics.

mesh_scattered: src/mesh.rs:89-99 (xorshift64 from 0x9E3779B97F4A7C15, 53-bit fraction).
tree_scattered: src/tree.rs:339-349 (xorshift64 from a seed, 24-bit fraction).
"""
import numpy as np

M64 = (1 << 64) - 1


def _xorshift_stream(state, count):
    out = []
    for _ in range(count):
        state ^= (state << 13) & M64
        state ^= state >> 7
        state ^= (state << 17) & M64
        out.append(state)
    return out


def mesh_scattered(count, box_size):
    """src/mesh.rs:89-99: (state >> 11) as f32 / 2^53 as f32 * box, in f32."""
    raw = _xorshift_stream(0x9E3779B97F4A7C15, 3 * count)
    f = np.float32
    vals = [f(f(r >> 11) / f(1 << 53)) * f(box_size) for r in raw]
    return np.array(vals, np.float32).reshape(count, 3)


def tree_scattered(count, box_size, seed):
    """src/tree.rs:339-349: (state >> 40) as f32 / 2^24 as f32 * box, in f32."""
    raw = _xorshift_stream(seed, 3 * count)
    f = np.float32
    vals = [f(f(r >> 40) / f(1 << 24)) * f(box_size) for r in raw]
    return np.array(vals, np.float32).reshape(count, 3)


def max_norm(v):
    return float(np.sqrt((np.asarray(v, np.float64) ** 2).sum(1)).max())


def max_diff(a, b):
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return float(np.sqrt((d ** 2).sum(1)).max())


def rel_errors(acc, ref):
    """Per-particle |a - a_ref| / |a_ref|."""
    acc = np.asarray(acc, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.sqrt(((acc - ref) ** 2).sum(1)) / np.maximum(np.sqrt((ref ** 2).sum(1)), 1e-300)


def lattice(n, box_size, centred=True, dtype=np.float32):
    h = box_size / n
    i = np.arange(n)
    q = (i + 0.5) * h if centred else i * h
    x, y, z = np.meshgrid(q, q, q, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(dtype)


# Cosmology of the reference (src/cosmology.rs:21-30).
OMEGA_M = 0.315


def plane_wave(n, box_size, amp, a_init, growth, growth_rate, e_of_a, dtype=np.float64):
    """src/integrator.rs:326-348: lattice q = i h, Psi_x = -(A/k) sin(k q_x), x = q + D Psi,
    p = a^2 f E D Psi. growth, growth_rate, e_of_a are callables of a."""
    k = 2 * np.pi / box_size
    h = box_size / n
    d = growth(a_init)
    f = growth_rate(a_init)
    e = e_of_a(a_init)
    i = np.arange(n)
    qx, qy, qz = np.meshgrid(i * h, i * h, i * h, indexing="ij")
    qx, qy, qz = qx.ravel(), qy.ravel(), qz.ravel()
    psi = -(amp / k) * np.sin(k * qx)
    x = np.mod(qx + d * psi, box_size)
    pos = np.stack([x, qy, qz], 1).astype(dtype)
    mom = np.stack([a_init * a_init * f * e * d * psi, 0 * qx, 0 * qx], 1).astype(dtype)
    return pos, mom


def mode_amplitude(delta, n, box_size):
    """src/integrator.rs:351-366: 2 mean(delta cos(k x)) over cell centres."""
    x = (np.arange(n) + 0.5) * box_size / n
    c = np.cos(2 * np.pi / box_size * x)
    return 2.0 * float((delta.reshape(n, n, n) * c[:, None, None]).sum()) / n ** 3


def fixed_point_delta32(pos32, n, box_size):
    """Density contrast of libtpm's fixed-point deposit with FP32 CIC weights and unit masses:
    s = x / h - 0.5 and the weights (wx wy) wz in FP32 as the reference computes them
    (src/mesh.rs:29-69), each corner's quantum round(w 2^32) with the last corner taking the
    remainder of round(2^32), summed exactly as integers; delta = c / mean - 1 in FP32."""
    pos32 = np.asarray(pos32, np.float32).reshape(-1, 3)
    h = np.float32(box_size) / np.float32(n)
    s = pos32 / h - np.float32(0.5)
    i0 = np.floor(s)
    base = i0.astype(np.int64)
    frac = (s - i0).astype(np.float32)
    total = np.int64(np.rint(2.0 ** 32))
    grid = np.zeros(n ** 3, np.int64)
    used = np.zeros(len(pos32), np.int64)
    one = np.float32(1)
    for c in range(8):
        d = ((c >> 2) & 1, (c >> 1) & 1, c & 1)
        w = [frac[:, a] if d[a] else one - frac[:, a] for a in range(3)]
        wt = (w[0] * w[1]) * w[2]
        q = np.rint(wt.astype(np.float64) * 2.0 ** 32).astype(np.int64) if c < 7 else total - used
        used += q
        idx = [(base[:, a] + d[a]) % n for a in range(3)]
        np.add.at(grid, (idx[0] * n + idx[1]) * n + idx[2], q)
    mean32 = np.float32(len(pos32)) / np.float32(n ** 3)
    return (grid.astype(np.float64) * 2.0 ** -32).astype(np.float32) / mean32 - np.float32(1)
