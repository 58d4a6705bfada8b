"""Pins the oracle's time-domain spectra (xpath.m:45-68,90-106 with the grid of
xpreferences.m:199-203): observables flagged in p.transforms{o}(1) are taken of
ifft(a(t)) * fsc over the stored step fields, fftshifted, at the points(1)
frequencies w = plotpts * dw around zero.

Checks: a linear SDE, whose interaction-picture steps are exact, against the
geometric series of its sampled exponential (a wrong transform sign, a
missing fsc / M, a shifted frequency index or a dropped first sample each
fail it); the fine run's two-half-step mean under checks; and the
Ornstein-Uhlenbeck spectrum of Examples/SDE_1D/Equilibrium.m:12-28 against its
own compare function p.ranges(1)/(pi (1 + w^2)), statistically."""
import math

import numpy as np

from oracle import xspde as X


def _linear_sde(L, pts, steps, checks):
    return dict(dimensions=1, points=[pts], ranges=[3.0], steps=[steps], checks=checks, noises=0,
                method="RK4", ensembles=[1, 1, 1], linear=lambda p: L,
                initial=lambda *a: 1.0 + 0j, observe=[lambda a, p: a, lambda a, p: a],
                transforms=[0, 1])


def _series(z, M, j):
    """sum_{n<M} z^n e^{2 pi i j n / M} summed in closed form."""
    q = z * np.exp(2j * np.pi * j / M)
    return (1 - q ** M) / (1 - q)


def test_spectrum_grid_matches_xpreferences():
    p = X.xpreferences(_linear_sde(-0.3 + 1.1j, 21, 3, 0))
    assert p.spectrum == 1
    dt = 3.0 / 20
    assert np.isclose(p.dw, 2 * np.pi / (3.0 + dt))
    assert np.isclose(p.fsc, math.sqrt(2 * np.pi) / p.dw)
    assert list(p.plotpts[:3]) == [-10, -9, -8] and p.plotpts[-1] == 10


def test_spectrum_linear_sde_geometric_series():
    L = -0.3 + 1.1j
    for pts, steps in ((21, 3), (20, 2)):          # odd and even numbers of samples
        p = X.xpreferences(_linear_sde(L, pts, steps, 0))
        _, av = X.run_path(p, 1, 1)
        M = pts * steps
        z = np.exp(L * p.dt / steps)               # one exact step
        j = p.plotpts
        expect = p.fsc / M * _series(z, M, j)
        assert np.allclose(av[1][0, :, 0], expect, rtol=1e-10, atol=1e-13)
        # the untransformed observable is the field at the stored points
        assert np.allclose(av[0][0, :, 0], z ** (steps * np.arange(pts)), rtol=1e-10)


def test_spectrum_fine_run_averages_half_steps():
    L = 0.2 - 0.7j
    pts, steps = 11, 2
    p = X.xpreferences(_linear_sde(L, pts, steps, 1))
    _, av = X.run_path(p, 1, 1)                    # fine run: two half steps per step
    M = pts * steps
    zf = np.exp(L * p.dt / (2 * steps))
    n = np.arange(1, M)
    samples = np.concatenate([[1.0], (zf ** (2 * n - 1) + zf ** (2 * n)) / 2])
    j = p.plotpts
    expect = p.fsc / M * (samples[None, :] * np.exp(2j * np.pi * np.outer(j, np.arange(M)) / M)).sum(1)
    assert np.allclose(av[1][0, :, 0], expect, rtol=1e-10, atol=1e-13)
    _, avc = X.run_path(p, 1, 2)                   # coarse run: whole steps
    zc = np.exp(L * p.dt / steps)
    assert np.allclose(avc[1][0, :, 0], p.fsc / M * _series(zc, M, j), rtol=1e-10, atol=1e-13)


def equilibrium(E=1200):
    """Examples/SDE_1D/Equilibrium.m:12-28 with the built-in forms' noise
    convention: da = (-a + w1 + i w2) dt, a(0) = (v1 + i v2)/sqrt(2)."""
    return dict(dimensions=1, points=[200], steps=[2], ranges=[25.0], seed=241, noises=2, inrandoms=2,
                ensembles=[E, 1, 1], method="MP",
                initial=lambda v, p: (v[0] + 1j * v[1]) / math.sqrt(2),
                deriv=lambda a, w, p: -a + w[0] + 1j * w[1],
                observe=[lambda a, p: np.abs(a) ** 2, lambda a, p: np.abs(a) ** 2],
                transforms=[0, 1])


def test_ou_spectrum_matches_equilibrium_compare():
    E = 1200
    p = X.xpreferences(equilibrium(E))
    _, av = X.run_path(p, 1, 1)
    w = p.w
    compare = p.ranges[0] / (np.pi * (1 + w ** 2))   # Equilibrium.m:28
    # stationary |a|^2 = 1 (Equilibrium.m:28, first compare)
    assert abs(np.mean(av[0][0, :, 0]) - 1.0) < 0.05
    S = av[1][0, :, 0].real
    sel = np.abs(w) < 3
    # each frequency's |a(w)|^2 is exponential-distributed over trajectories:
    # relative standard error 1/sqrt(E) per point; the mean over the band
    rel = S[sel] / compare[sel] - 1
    assert abs(rel.mean()) < 4 / math.sqrt(E * sel.sum())
    assert np.sqrt(np.mean(rel ** 2)) < 3 / math.sqrt(E)
