"""The GPU Morton sort (hand-written LSD radix passes) against numpy's stable argsort of the same
keys: the octree order of Tree::build (src/tree.rs:215-236) for every sorted depth, 32-bit keys
(depth <= 10) and 64-bit keys, ragged sizes, tile boundaries and heavy key duplication."""
import numpy as np
import pytest

from helpers import OMEGA_M
from morton import octree_order

pytestmark = pytest.mark.gpu

BOX = 100.0


def _check(treepm, pos, depth, mesh=16):
    n = len(pos)
    print(f"[gpu] morton order n={n} depth={depth}", flush=True)
    with treepm.Solver(mesh, BOX, OMEGA_M, max_particles=max(n, 1)) as s:
        s.set_tuning(16, 0, fft_path=0, tree_depth=depth, small_node_direct=0)
        order, keys = s.morton_order(pos)
    ref_order, ref_keys = octree_order(pos, BOX, depth)
    if depth <= 10:
        # 32-bit sort keys carry octree levels 1-10: the levels below are cleared
        ref_keys = ref_keys & ~np.uint64((1 << 33) - 1)
    assert np.array_equal(order, ref_order), depth
    assert np.array_equal(keys, ref_keys), depth


@pytest.mark.parametrize("depth", [8, 10, 12, 21])
def test_uniform_ragged(treepm, depth):
    rng = np.random.default_rng(depth)
    _check(treepm, rng.random((100_003, 3)) * BOX, depth)


@pytest.mark.parametrize("n", [1, 2, 63, 4095, 4096, 4097, 3 * 4096 + 17])
def test_tile_boundaries(treepm, n):
    rng = np.random.default_rng(n)
    _check(treepm, rng.random((n, 3)) * BOX, 21)
    _check(treepm, rng.random((n, 3)) * BOX, 8)


def test_duplicates_keep_input_order(treepm):
    rng = np.random.default_rng(7)
    pos = np.repeat(rng.random((37, 3)) * BOX, 3000, axis=0)  # 111,000 particles on 37 points
    pos = pos[rng.permutation(len(pos))]
    _check(treepm, pos, 21)
    _check(treepm, np.full((20_000, 3), 33.3), 10)


def test_clustered_and_lattice(treepm):
    with treepm.Solver(64, BOX, OMEGA_M, max_particles=300_000) as ic:
        nfw = ic.nfw_halos(treepm.NfwSpec.clustered(seed=2, n_halos=500), 300_000)
        lat, _ = ic.zeldovich(1, 2.0, 0.3 * BOX / 64)
    for depth in (8, 21):
        _check(treepm, nfw, depth)
        _check(treepm, lat, depth, mesh=64)
