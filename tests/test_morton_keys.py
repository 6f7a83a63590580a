"""The numpy Morton-key checker against the reference's octant rule (src/tree.rs:216-218) and
closed forms, so the GPU sort is compared with a pinned order."""
import numpy as np

from morton import KEY_BITS, morton_keys, octree_order


def test_level_one_digit_is_the_reference_octant():
    rng = np.random.default_rng(3)
    box = 100.0
    pos = rng.random((5000, 3)) * box
    digit = morton_keys(pos, box) >> np.uint64(3 * (KEY_BITS - 1))
    c = box / 2  # the root's centre (src/tree.rs:180-181)
    octant = (pos[:, 0] >= c).astype(np.uint64) | ((pos[:, 1] >= c).astype(np.uint64) << np.uint64(1)) | (
        (pos[:, 2] >= c).astype(np.uint64) << np.uint64(2))
    assert np.array_equal(digit, octant)


def test_deeper_digits_follow_the_child_cubes():
    # Level l digit = octant of the particle within its level-(l-1) cube: compare with a direct
    # recursive descent (centre +- quarter, src/tree.rs:243-245) in FP64.
    rng = np.random.default_rng(4)
    box = 1.0
    pos = rng.random((300, 3)) * box
    keys = morton_keys(pos, box)
    for i in range(len(pos)):
        centre = np.full(3, box / 2)
        half = box / 2
        for level in range(1, 12):
            o = int(pos[i, 0] >= centre[0]) | int(pos[i, 1] >= centre[1]) << 1 | int(pos[i, 2] >= centre[2]) << 2
            d = int(keys[i] >> np.uint64(3 * (KEY_BITS - level))) & 7
            assert d == o, (i, level)
            half /= 2
            centre = centre + np.where([o & 1, o & 2, o & 4], half, -half)


def test_axis_keys_and_clamping():
    box = 8.0
    # a particle in cell i along x only: the key is i's bits spread to every third position
    i = 0b101101
    x = (i + 0.5) * box / (1 << KEY_BITS)
    k = int(morton_keys(np.array([[x, 0.0, 0.0]]), box)[0])
    assert k == sum(((i >> b) & 1) << (3 * b) for b in range(KEY_BITS))
    # the box face and beyond clamp to the last cell; negatives to the first
    k = morton_keys(np.array([[box, box, box], [-1.0, -1.0, -1.0]]), box)
    assert int(k[0]) == (1 << (3 * KEY_BITS)) - 1 and int(k[1]) == 0


def test_order_is_stable_below_the_depth():
    pos = np.array([[0.9, 0.9, 0.9], [0.1, 0.1, 0.1], [0.91, 0.9, 0.9], [0.11, 0.1, 0.1]])
    order, _ = octree_order(pos, 1.0, 1)
    assert list(order) == [1, 3, 0, 2]
