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
