"""Writes the committed golden vectors under tests/golden/.

Each vector is computed here by an implementation independent of the oracle and the HIP
library (numpy FP64 transforms, scipy's erfc, direct numpy pair sums), from the reference
tests' own deterministic inputs (src/mesh.rs:89-99, src/tree.rs:339-349):

  cic_5000.npz      CIC deposit of the 5000-particle mesh.rs input on 16^3, L = 1 (FP64)
  bruteforce.npz    short-range brute-force sums for the tree.rs inputs (400 @ seed 7,
                    r_cut 2; 3000 @ seed 11, r_cut 2.5), L = 10, r_s = 0.5, eps = 0.05 (FP64)
  fft.npz           numpy FFTs of seeded complex lines (n = 6, 8, 16, 47, 94) and of a 12^3 grid
  srfactor.npz      S(r) = erfc(u) + 2u/sqrt(pi) e^{-u^2} on a grid of r for r_s = 1 (scipy)

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
from scipy.special import erfc

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from helpers import mesh_scattered, tree_scattered  # noqa: E402


def cic_fp64(pos, n, box, mass):
    pos = np.asarray(pos, np.float64)
    h = box / n
    s = pos / h - 0.5
    i0 = np.floor(s)
    f = s - i0
    i0 = i0.astype(np.int64)
    grid = np.zeros(n ** 3)
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = np.where(dx, f[:, 0], 1 - f[:, 0]) * np.where(dy, f[:, 1], 1 - f[:, 1]) * \
                    np.where(dz, f[:, 2], 1 - f[:, 2])
                ix = (i0[:, 0] + dx) % n
                iy = (i0[:, 1] + dy) % n
                iz = (i0[:, 2] + dz) % n
                np.add.at(grid, (ix * n + iy) * n + iz, mass * w)
    return grid


def brute_fp64(pos, box, r_split, r_cut, eps, k):
    pos = np.asarray(pos, np.float64)
    d = pos[:, None, :] - pos[None, :, :]
    d -= box * np.round(d / box)
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    mask = r2 <= r_cut * r_cut
    r = np.sqrt(np.where(mask, r2, 1.0))
    u = r / (2 * r_split)
    S = erfc(u) + 2 / np.sqrt(np.pi) * u * np.exp(-u * u)
    s = np.where(mask, r2, 1.0) + eps * eps
    f = np.where(mask, -k * S / (s * np.sqrt(s)), 0.0)
    return (d * f[..., None]).sum(1)


def main():
    pos = mesh_scattered(5000, 1.0)
    np.savez_compressed(os.path.join(HERE, "cic_5000.npz"), pos=pos, grid=cic_fp64(pos, 16, 1.0, 0.25))

    p400 = tree_scattered(400, 10.0, 7)
    p3000 = tree_scattered(3000, 10.0, 11)
    np.savez_compressed(os.path.join(HERE, "bruteforce.npz"),
                        pos400=p400, acc400=brute_fp64(p400, 10.0, 0.5, 2.0, 0.05, 1.0),
                        pos3000=p3000, acc3000=brute_fp64(p3000, 10.0, 0.5, 2.5, 0.05, 1.0))

    rng = np.random.default_rng(20240601)
    out = {}
    for n in (6, 8, 16, 47, 94):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"x{n}"] = x
        out[f"f{n}"] = np.fft.fft(x)
    g = rng.standard_normal((12, 12, 12))
    out["g12"] = g
    out["G12"] = np.fft.fftn(g)
    np.savez_compressed(os.path.join(HERE, "fft.npz"), **out)

    r = np.linspace(0.0, 10.0, 401)
    u = r / 2.0
    np.savez_compressed(os.path.join(HERE, "srfactor.npz"), r=r,
                        S=erfc(u) + 2 / np.sqrt(np.pi) * u * np.exp(-u * u))


if __name__ == "__main__":
    main()
