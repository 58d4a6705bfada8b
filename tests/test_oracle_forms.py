"""Pins the oracle's observe and initial forms that have no reference known
answer: the equal-time correlation 'corr', the low-mode occupation 'lowk' and
the Wigner-vacuum 'coherent' initial state.

* corr against the O(Ns^2) double sum (1/Ns) sum_x a*(x) a(x + r) on lattices
  of unequal and odd sides, where a flipped r or a 1/Ns^2 normalisation gives a
  different result (both are checked to fail);
* lowk against a direct DFT sum over the chosen modes;
* the coherent form on vacuum samples: per-mode occupation 1/2 for every k != 0
  (variance 1/2 per mode, Applications/Phase-space/GBSObserve/Np.m:11-16
  symmetric ordering), zero mean of lowk, and |alpha|^2 Ns dv + 1/2 at k = 0.
"""
import math

import numpy as np
import pytest

from oracle import forms
from oracle import xspde as X


def _lattice(nx, ny, dx=1.0, dy=1.0, E=3):
    """A small 2D project (or 1D with ny = 0) in the oracle's parameter form."""
    if ny:
        p = dict(dimensions=3, points=[2, nx, ny], ranges=[1.0, dx * (nx - 1), dy * (ny - 1)],
                 ensembles=[E, 1, 1], noises=0)
    else:
        p = dict(dimensions=2, points=[2, nx], ranges=[1.0, dx * (nx - 1)], ensembles=[E, 1, 1], noises=0)
    return X.xpreferences(p)


def _field(q, seed):
    rng = np.random.default_rng(seed)
    shape = (1,) + q.space + (q.ensembles[0],)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _corr_brute(a, sign=+1, norm=None):
    """(1/Ns) sum_x conj(a(x)) a(x + sign r) by explicit loops, a[1, x, (y), E]."""
    space = a.shape[1:-1]
    ns = int(np.prod(space))
    norm = ns if norm is None else norm
    out = np.zeros(a.shape, complex)
    for r in np.ndindex(*space):
        acc = np.zeros(a.shape[-1], complex)
        for x in np.ndindex(*space):
            xr = tuple((xi + sign * ri) % n for xi, ri, n in zip(x, r, space))
            acc += np.conj(a[(0,) + x]) * a[(0,) + xr]
        out[(0,) + r] = acc / norm
    return out


@pytest.mark.parametrize("nx,ny", [(4, 4), (5, 3), (6, 0)])
def test_corr_matches_double_sum(nx, ny):
    q = _lattice(nx, ny)
    a = _field(q, 1 + nx)
    got = forms.make_observe("corr")(a, q)
    ref = _corr_brute(a)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    # the test discriminates: a flipped r and a 1/Ns^2 normalisation both fail
    assert np.abs(got - _corr_brute(a, sign=-1)).max() > 1e-3 * np.abs(ref).max()
    ns = q.nspace
    assert np.abs(got - _corr_brute(a, norm=ns * ns)).max() > 1e-3 * np.abs(ref).max()
    # r = 0 is the lattice mean of |a|^2
    assert np.allclose(got[(0,) + (0,) * len(q.space)], np.mean(np.abs(a[0]) ** 2, axis=tuple(range(len(q.space)))))


def _lowk_brute(a, q, modes):
    """sum over the modes k != 0 of |a_k|^2 - 1/2, a_k = sqrt(dv/Ns) sum_x a(x) e^{-i k.x}."""
    space = q.space if len(q.space) == 2 else q.space + (1,)
    nx, ny = space
    E = a.shape[-1]
    af = a[0].reshape(nx, ny, E)
    tot = np.zeros(E)
    for kx, ky in modes:
        if (kx, ky) == (0, 0):
            continue
        ak = np.zeros(E, complex)
        for x in range(nx):
            for y in range(ny):
                ak += af[x, y] * np.exp(-2j * np.pi * (kx * x / nx + ky * y / ny))
        ak *= math.sqrt(q.dv / q.nspace)
        tot += np.abs(ak) ** 2 - 0.5
    return tot


@pytest.mark.parametrize("nx,ny,dx,dy", [(4, 4, 1.0, 1.0), (6, 5, 0.5, 2.0)])
def test_lowk_matches_direct_dft(nx, ny, dx, dy):
    q = _lattice(nx, ny, dx, dy)
    a = _field(q, 7)
    modes = [(0, 0), (1, 0), (0, 1), (nx - 1, 0), (2, ny - 1)]
    got = np.asarray(forms.make_observe("lowk", {"modes": modes})(a, q)).reshape(-1)
    ref = _lowk_brute(a, q, modes)
    assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
    # a count selects the lowest modes: the first entry is k = 0, which is skipped
    lm = forms.lowest_modes(nx, ny, 5)
    assert lm[0] == (0, 0)
    got5 = np.asarray(forms.make_observe("lowk", {"modes": 5})(a, q)).reshape(-1)
    assert np.allclose(got5, _lowk_brute(a, q, lm), rtol=1e-12, atol=1e-12)


def test_coherent_vacuum_occupation():
    """Wigner vacuum of the coherent form: every mode k != 0 holds 1/2 on average,
    so lowk has mean 0; k = 0 holds |alpha|^2 Ns dv + 1/2.  (A variance 1 per
    mode, or 1/4, fails these bounds by many standard errors.)"""
    E = 4000
    alpha = 1.5 - 0.5j
    p = dict(dimensions=3, points=[2, 8, 6], ranges=[1.0, 3.5, 10.0], ensembles=[E, 1, 1], noises=0,
             inrandoms=2, gpu=dict(initial="coherent", alpha=alpha, observe=["abs2_ave"], modes=6))
    q = X.xpreferences(forms.apply_forms(p))
    v = X.xrandom(q, 3)
    a = X.initial_field(q, v)
    nx, ny = q.space
    A = np.fft.fftn(a[0], axes=(0, 1)) * math.sqrt(q.dv / q.nspace)      # [kx, ky, E]
    occ = np.abs(A) ** 2
    n0 = abs(alpha) ** 2 * q.nspace * q.dv
    mean0 = occ[0, 0].mean()
    # |a_0|^2 = |c + z|^2 with z of variance 1/2: mean n0 + 1/2, sd ~ sqrt(2 n0 * 1/2)
    assert abs(mean0 - (n0 + 0.5)) < 5 * math.sqrt(n0 + 0.25) / math.sqrt(E)
    rest = occ.reshape(nx * ny, E)[1:]
    # per mode |a_k|^2 is exponential with mean 1/2 (sd 1/2)
    per_mode = rest.mean(axis=1)
    assert np.all(np.abs(per_mode - 0.5) < 5 * 0.5 / math.sqrt(E))
    assert abs(rest.mean() - 0.5) < 5 * 0.5 / math.sqrt(E * rest.shape[0])
    # real and imaginary parts of a_k each have variance 1/4 (1/2 per complex mode)
    assert abs(np.var(A.reshape(nx * ny, E)[1:].real) - 0.25) < 0.01
    lowk = np.asarray(forms.make_observe("lowk", p["gpu"])(a, q)).reshape(-1)
    nm = len([m for m in forms.lowest_modes(nx, ny, 6) if m != (0, 0)])
    assert abs(lowk.mean()) < 5 * 0.5 * math.sqrt(nm) / math.sqrt(E)
    # per-site: a - alpha has variance 1/(2 dv) (1/(4 dv) per real component)
    dev = a[0] - alpha
    assert abs(np.var(dev.real) * q.dv - 0.25) < 0.01
    assert abs(np.var(dev.imag) * q.dv - 0.25) < 0.01
