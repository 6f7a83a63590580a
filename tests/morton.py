"""Morton keys and the octree order in numpy: the checker of the GPU sort.

The octree's child octant of a particle is bit 0 = x >= centre, bit 1 = y, bit 2 = z
(src/tree.rs:216-218); the 63-bit key interleaves floor(x / L * 2^21) per axis so that the 3-bit
digit of level l (l = 1 for the root's children) is that octant at level l.
"""
import numpy as np

KEY_BITS = 21


def cell_coords(pos, box):
    q = np.floor((np.asarray(pos, np.float64) / box) * float(1 << KEY_BITS))
    return np.clip(q, 0, (1 << KEY_BITS) - 1).astype(np.uint64)


def morton_keys(pos, box):
    q = cell_coords(pos, box)
    key = np.zeros(len(q), np.uint64)
    one = np.uint64(1)
    for b in range(KEY_BITS):
        for axis in range(3):
            bit = (q[:, axis] >> np.uint64(b)) & one
            key |= bit << np.uint64(3 * b + axis)
    return key


def octree_order(pos, box, depth):
    """Stable order by the key digits of levels 1..depth (input order below)."""
    keys = morton_keys(pos, box)
    top = keys >> np.uint64(3 * (KEY_BITS - depth))
    order = np.argsort(top, kind="stable").astype(np.uint32)
    return order, keys[order]
