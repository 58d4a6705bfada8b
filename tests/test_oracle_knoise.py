"""Pins the oracle's k-space filtered noises and randoms (xnoise.m:39-50,
xrandom.m:24-29, xift.m): for real N(0,1) values z_k at every k site (FFT
order) and a filter F, W(x) = c xift(F z)(x) has the closed-form covariances

    <W(x) W*(x')> = c^2 / N^2 sum_k |F(k)|^2 exp(+2 pi i k.(x - x') / N)
    <W(x) W(x')>  = c^2 / N^2 sum_k F(k)^2   exp(+2 pi i k.(x + x') / N)

with c^2 = nscale * knoisefactor (noises) or knoisefactor (randoms),
knoisefactor = N / dv (xpreferences.m:142).  The filters are not symmetric in
k, so a wrong transform direction, axis or scale fails."""
import numpy as np
import pytest

from oracle import xspde as X


def _theory(F, c2, shape):
    N = int(np.prod(shape))
    ks = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    sites = list(np.ndindex(*shape))
    cov = np.zeros((N, N), complex)
    rel = np.zeros((N, N), complex)
    for i, x in enumerate(sites):
        for j, y in enumerate(sites):
            ph_m = sum(k * (a - b) / n for k, a, b, n in zip(ks, x, y, shape))
            ph_p = sum(k * (a + b) / n for k, a, b, n in zip(ks, x, y, shape))
            cov[i, j] = c2 / N ** 2 * np.sum(np.abs(F) ** 2 * np.exp(2j * np.pi * ph_m))
            rel[i, j] = c2 / N ** 2 * np.sum(F ** 2 * np.exp(2j * np.pi * ph_p))
    return cov, rel


def _filter(p):
    k = p.kx if p.dimensions == 2 else p.kx + 0.7 * p.ky
    return np.exp(-(k - 0.4) ** 2) * (1 + 0.5j * np.sin(k))


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("which", ["noise", "random"])
def test_filtered_noise_covariance(dims, which):
    shape = (8,) if dims == 2 else (4, 4)
    E = 40000
    p = dict(dimensions=dims, points=[3] + list(shape), ranges=[1.0] + [4.0] * (dims - 1), steps=[2] + [1] * (dims - 1),
             ensembles=[E, 1, 1], noises=0, inrandoms=0)
    if which == "noise":
        p.update(knoises=1, nfilter=lambda w, p: w * _filter(p))
    else:
        p.update(krandoms=1, rfilter=lambda w, p: w * _filter(p))
    q = X.xpreferences(p)
    q.dtr = q.dt / q.steps[0]
    if which == "noise":
        w = X.noise_draw(q, 1, 0)
        c2 = q.nscale * q.knoisefactor
    else:
        w = X.xrandom(q, 1)[0]
        c2 = q.knoisefactor
    assert w.shape == (1,) + shape + (E,)
    W = w[0].reshape(-1, E, order="C")
    F = _filter(q)[0, ..., 0] if dims == 3 else _filter(q).reshape(-1)[:shape[0]]
    F = np.asarray(F).reshape(shape)
    cov, rel = _theory(F, c2, shape)
    emp_cov = W @ W.conj().T / E
    emp_rel = W @ W.T / E
    scale = np.abs(np.diag(cov)).max()
    assert np.abs(emp_cov - cov).max() < 0.05 * scale
    assert np.abs(emp_rel - rel).max() < 0.05 * scale
    # a plausible mistake: forward instead of inverse transform
    cov_wrong = np.conj(cov)
    assert np.abs(emp_cov - cov_wrong).max() > 0.1 * scale


def test_identity_filter_keeps_xspace_variance():
    """Default filter (xnfilter: identity): each site's k-noise has
    <|W|^2> = nscale / dv, the variance of one x-space noise component."""
    p = dict(dimensions=2, points=[3, 16], ranges=[1.0, 5.0], steps=[4, 1], ensembles=[20000, 1, 1], knoises=1)
    q = X.xpreferences(p)
    assert q.noises == 0 and q.inrandoms == 0
    w = X.noise_draw(q, 3, 1)
    var = np.mean(np.abs(w) ** 2)
    assert abs(var / (q.nscale / q.dv) - 1) < 0.03
