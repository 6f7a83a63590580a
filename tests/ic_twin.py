"""numpy restatement of libtpm's Zel'dovich generator (test infrastructure): the same
counter-based noise (SplitMix64 finaliser + Box-Muller, rounded to float32), the power law
P(k) = k^-2 exp(-k/k_c), Psi_k = i k F_k / k^2 with the reference's wavenumber convention
(src/fft.rs:86-95, src/ics.rs:134-154), rescaled to a 3D displacement RMS, x = q + Psi on the
lattice q = i h (src/ics.rs:167) wrapped into [0, L)."""
import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def noise(n, seed):
    i = np.arange(n ** 3, dtype=np.uint64)
    x = splitmix64(np.uint64(seed) ^ splitmix64(i))
    y = splitmix64(x)
    u1 = ((x >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = (y >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return (np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)).astype(np.float32)


def wavenumbers(n, box):
    i = np.arange(n)
    k = np.where(i <= n // 2, i, i - n) * (2 * np.pi / box)
    if n % 2 == 0:
        k[n // 2] = 0.0
    return k


def zeldovich(n, box, seed, k_c, disp_rms, mom_factor=0.0):
    w = noise(n, seed).reshape(n, n, n).astype(np.float64)
    W = np.fft.rfftn(w)
    k = wavenumbers(n, box)
    kx = k[:, None, None]
    ky = k[None, :, None]
    kz = k[None, None, : n // 2 + 1]
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        kk = np.sqrt(k2)
        G = np.where(k2 > 0, -np.sqrt(n ** 3 / box ** 3) * np.sqrt(np.exp(-kk / k_c) / k2) / k2, 0.0)
    psi = [np.fft.irfftn(G * (-1j * ka) * W, s=(n, n, n), axes=(0, 1, 2)) for ka in (kx, ky, kz)]
    rms = np.sqrt(sum((p ** 2).sum() for p in psi) / n ** 3)
    scale = disp_rms / rms
    h = box / n
    q = np.arange(n) * h
    qs = np.meshgrid(q, q, q, indexing="ij")
    pos = np.stack([np.mod(qs[a] + scale * psi[a], box) for a in range(3)], -1).reshape(-1, 3)
    mom = np.stack([mom_factor * scale * psi[a] for a in range(3)], -1).reshape(-1, 3)
    return pos, mom


def bbks_transfer(k, gamma):
    """BBKS transfer (src/cosmology.rs:93-103), q = k / Gamma."""
    q = np.asarray(k, np.float64) / gamma
    with np.errstate(divide="ignore", invalid="ignore"):
        poly = 1.0 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4
        t = np.log1p(2.34 * q) / (2.34 * q) * poly ** -0.25
    return np.where(q == 0, 1.0, t)


def zeldovich_bbks(n, box, seed, amplitude, n_s, gamma, d, mom_factor):
    """Lattice Zel'dovich ICs from P(k) = A k^n_s T^2 scaled by D (src/ics.rs:78-182), the same noise."""
    w = noise(n, seed).reshape(n, n, n).astype(np.float64)
    W = np.fft.rfftn(w)
    k = wavenumbers(n, box)
    kx = k[:, None, None]
    ky = k[None, :, None]
    kz = k[None, None, : n // 2 + 1]
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        kk = np.sqrt(k2)
        p = amplitude * kk ** n_s * bbks_transfer(kk, gamma) ** 2
        G = np.where(k2 > 0, -np.sqrt(n ** 3 / box ** 3) * np.sqrt(p) / k2, 0.0)
    psi = [np.fft.irfftn(G * (-1j * ka) * W, s=(n, n, n), axes=(0, 1, 2)) for ka in (kx, ky, kz)]
    h = box / n
    q = np.arange(n) * h
    qs = np.meshgrid(q, q, q, indexing="ij")
    pos = np.stack([np.mod(qs[a] + d * psi[a], box) for a in range(3)], -1).reshape(-1, 3)
    mom = np.stack([mom_factor * d * psi[a] for a in range(3)], -1).reshape(-1, 3)
    return pos, mom
