"""Pins the oracle's counter-based Gaussian stream (oracle/philox.py)."""
import numpy as np
import pytest
from scipy import stats

from oracle import philox
from oracle import xspde as X

# Random123 known-answer vectors for philox4x32_10 (ctr, key) -> out
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    r = philox.philox4x32_10(*[np.uint64(c) for c in ctr], *[np.uint64(k) for k in key])
    assert tuple(int(x) for x in r) == out


def test_uniform_open_interval():
    x = np.array([0, 1, 2 ** 31, 2 ** 32 - 1], dtype=np.uint64)
    u = philox.uniform_open01(x)
    assert np.all(u > 0) and np.all(u < 1)
    assert u[0] == 2.0 ** -33


def test_normals_moments_and_ks():
    n = philox.normals(7, philox.TAG_NOISE, 3, np.arange(64), 4096)
    z = n.ravel()
    assert abs(z.mean()) < 5 * (1 / np.sqrt(z.size))
    assert abs(z.var() - 1) < 5 * np.sqrt(2 / z.size)
    assert stats.kstest(z, "norm").pvalue > 1e-4


def test_streams_distinct():
    a = philox.normals(1, philox.TAG_NOISE, 0, np.arange(4), 16)
    b = philox.normals(2, philox.TAG_NOISE, 0, np.arange(4), 16)
    c = philox.normals(1, philox.TAG_NOISE, 1, np.arange(4), 16)
    d = philox.normals(1, philox.TAG_GAUSS_INIT, 0, np.arange(4), 16)
    for other in (b, c, d):
        assert np.abs(a - other).max() > 0.1
    # same counter, different slicing of trajectories -> identical values
    e = philox.normals(1, philox.TAG_NOISE, 0, np.arange(2, 4), 16)
    np.testing.assert_array_equal(a[2:], e)


def test_noise_scale_is_one_over_dtr_dv():
    # xnoise.m:31-32 with nscale = steps*(1+checks)/dt, noisefactor = 1/dv
    p = X.xpreferences(dict(dimensions=2, points=[11, 64], ranges=[1.0, 8.0], steps=[4, 1],
                            noises=2, ensembles=[256, 1, 1]))
    w = X.noise_draw(p, 1, 0)
    dtr_fine = p.dt / (2 * p.steps[0])
    var = np.var(w)
    assert w.shape == (2, 64, 256)
    assert abs(var * dtr_fine * p.dv - 1) < 0.03


def test_coarse_noise_is_pair_average():
    # xnoise.m:33-34: the coarse step uses (w_{2m} + w_{2m+1})/2
    p = dict(dimensions=1, points=[3], ranges=[1.0], steps=[1], checks=1, noises=1,
             ensembles=[5, 1, 1], method="MP",
             deriv=lambda a, w, p: w[0:1] + 0 * a, initial=lambda v, p: 0 * v)
    p = X.xpreferences(p)
    seen = {"fine": [], "coarse": []}

    def run(nc):
        got = []

        def fn(q):
            w = X.noise_draw(p, 1, q)
            got.append((q, w))
            return w
        a, _ = X.run_path(p, 1, nc, noise_fn=fn)
        return a, got
    af, gf = run(1)
    ac, gc = run(2)
    assert [q for q, _ in gf] == [q for q, _ in gc] == [0, 1, 2, 3]
    # additive noise: the endpoint is the sum of dt * w for fine, 2dt * mean for coarse
    wsum = sum(w for _, w in gf)
    np.testing.assert_allclose(af, wsum * p.dt / 2, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ac, wsum * p.dt / 2, rtol=1e-12, atol=1e-12)
