"""Clustered NFW workload generator (tpm_nfw_halos): determinism, the halo catalogue against a
numpy restatement, and the radial profile of the largest halo against the NFW enclosed mass."""
import numpy as np
import pytest

from ic_twin import splitmix64

pytestmark = pytest.mark.gpu


def _u01(seed, idx, stream):
    x = splitmix64(np.uint64(seed) ^ splitmix64(np.asarray(idx, np.uint64) * np.uint64(4) + np.uint64(stream)))
    return ((x >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53


def catalogue(seed, nh, n, frac=0.7, mmin=1e11, mmax=1e15, slope=1.9, cmin=5.0, cmax=15.0, box=100.0):
    hseed = int(splitmix64(np.uint64(seed) ^ np.uint64(0xA5A5A5A5A5A5A5A5)))
    h = np.arange(nh, dtype=np.uint64)
    e = 1.0 - slope
    a0, a1 = mmin ** e, mmax ** e
    mass = (a0 + _u01(hseed, h, 0) * (a1 - a0)) ** (1.0 / e)
    nhp = int(round(frac * n))
    want = mass / mass.sum() * nhp
    cnt = np.floor(want).astype(np.int64)
    order = sorted(range(nh), key=lambda k: (-(want[k] - np.floor(want[k])), k))
    k = 0
    while cnt.sum() < nhp:
        cnt[order[k % nh]] += 1
        k += 1
    centre = np.stack([box * _u01(hseed, h, s) for s in (1, 2, 3)], 1)
    r200 = np.cbrt(3.0 * mass / (4.0 * np.pi * 200.0 * 2.77536627e11))
    conc = cmin + (cmax - cmin) * _u01(hseed + 1, h, 0)
    first = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    return mass, cnt, centre, r200, conc, first, nhp


def test_nfw_generator(treepm):
    n, nh = 200000, 200
    s = treepm.Solver(16, 100.0, max_particles=n)
    spec = treepm.NfwSpec.clustered(seed=9, n_halos=nh)
    print(f"[gpu] nfw n={n} halos={nh}", flush=True)
    p1 = s.nfw_halos(spec, n)
    p2 = s.nfw_halos(spec, n)
    assert np.array_equal(p1, p2)
    assert np.all((p1 >= 0) & (p1 < 100.0))
    mass, cnt, centre, r200, conc, first, nhp = catalogue(9, nh, n)
    assert cnt.sum() == nhp == 140000
    big = int(np.argmax(cnt))
    pts = p1[first[big]:first[big] + cnt[big]]
    d = pts - centre[big]
    d -= 100.0 * np.round(d / 100.0)
    r = np.sqrt((d ** 2).sum(1))
    assert r.max() <= r200[big] * (1 + 1e-9)
    m = lambda x: np.log1p(x) - x / (1 + x)
    c = conc[big]
    for frac_r in (0.1, 0.3, 0.6):
        expect = m(c * frac_r) / m(c)
        got = np.mean(r < frac_r * r200[big])
        sigma = np.sqrt(expect * (1 - expect) / len(r))
        assert abs(got - expect) < 5 * sigma + 1e-3
    # isotropy: mean direction near zero
    assert np.abs((d / r[:, None]).mean(0)).max() < 5 / np.sqrt(len(r))
