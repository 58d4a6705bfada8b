"""The host's ensemble bookkeeping and statistics (xspde_matlab_amd.xsim:
ensemble_moments, statistics) on synthetic path averages, without a GPU:

* against the oracle's xensemble (Simulator/xensemble.m:84-105) and xsim_stats
  (Simulator/xsim.m:94-108,144-188) fed the same averages, every column (mean,
  step error, sampling error) and the error vector, for order 0 and 2, checks
  0 and 1, several sub-ensembles and parallel ensembles;
* against closed forms: the sampling error is the standard error of the mean of
  the e2*e3 sub-ensemble values (scipy.stats.sem), and with order-2 Richardson
  extrapolation of D = exact + c h^2 the mean is exact and the step error is
  c (h/2)^2.
"""
import types

import numpy as np
import pytest
from scipy import stats

from oracle import xspde as X
from xspde_matlab_amd.xsim import ensemble_moments, preferences, split_averages, statistics


def _info(q, ns_list):
    """A stand-in for xsh_info: offsets and site counts of the observables."""
    off, offs = 0, []
    for ns in ns_list:
        offs.append(off)
        off += q["points"][0] * ns
    return types.SimpleNamespace(obs_offset=offs, obs_sites=list(ns_list), avlen=off)


def _project(e2, e3, checks, order, nx=5, observe=("a", "abs2_ave")):
    return dict(name="synthetic", dimensions=2, points=[7, nx], ranges=[1.0, 4.0], steps=[2, 1], checks=checks,
                order=order, ensembles=[4, e2, e3], noises=0,
                gpu=dict(deriv="zero", initial="const", observe=list(observe)))


def _synthetic_averages(q, ns_list, e2, nchk, seed):
    """av[npar][j, nc] flat complex (MATLAB order per observable: time fastest)
    and the same averages as the oracle's [lines, points, sites] arrays."""
    rng = np.random.default_rng(seed)
    npts = q["points"][0]
    flat, arrays = [], []
    for _ in range(q["ensembles"][2]):
        f = np.zeros((e2, nchk, sum(npts * n for n in ns_list)), complex)
        a = {}
        for j in range(e2):
            for nc in range(nchk):
                parts, avs = [], []
                for ns in ns_list:
                    av = rng.standard_normal((1, npts, ns)) + 1j * rng.standard_normal((1, npts, ns))
                    av += 1.0 + 0.5 * nc              # fine and coarse differ a little
                    avs.append(av)
                    parts.append(av.reshape(-1, order="F"))
                f[j, nc] = np.concatenate(parts)
                a[(j + 1, nc + 1)] = avs
        flat.append(f)
        arrays.append(a)
    return flat, arrays


OUTPUTS = [None, lambda av, p: np.abs(av[1]) ** 2]


@pytest.mark.parametrize("e2,e3,checks,order", [(3, 2, 1, 0), (4, 1, 1, 2), (1, 3, 1, 1), (5, 1, 0, 0)])
def test_host_statistics_match_oracle(e2, e3, checks, order):
    p = _project(e2, e3, checks, order)
    q = preferences(p)
    nx = q["points"][1]
    ns_list = [nx, 1]
    info = _info(q, ns_list)
    outs = [lambda av, p: av[0], OUTPUTS[1]]
    nchk = 1 + checks
    flat, arrays = _synthetic_averages(q, ns_list, e2, nchk, seed=e2 * 10 + order)
    # host
    data = None
    for npar in range(e3):
        dp = ensemble_moments(q, info, outs, flat[npar])
        data = dp if data is None else [a + b for a, b in zip(data, dp)]
    err_h, res_h = statistics(q, data)
    # oracle, fed the same averages through its own bookkeeping
    po = X.xpreferences(dict(p, observe=[lambda a, p: a[0:1], lambda a, p: np.mean(a[0:1], axis=1)],
                             output=[None, OUTPUTS[1]]))
    dref = None
    for npar in range(1, e3 + 1):
        dp = X.xensemble(npar, po, path_fn=lambda p_, j, nc, n_: arrays[n_ - 1][(j, nc)])
        dref = dp if dref is None else [a + b for a, b in zip(dref, dp)]
    for o in range(2):
        assert np.allclose(data[o], dref[o], rtol=1e-13, atol=1e-13), "D, D^2 sums o=%d" % o
    po._avshape = [(1, q["points"][0], nx), (1, q["points"][0])]
    err_r, res_r = X.xsim_stats(po, dref)
    sx = max(3, 1 + nchk)
    for o in range(2):
        r = res_r[o].reshape(-1, res_r[o].shape[-1])
        assert np.allclose(res_h[o][:, 0], r[:, 0], rtol=1e-13, atol=1e-14)
        assert np.allclose(res_h[o][:, 2], r[:, sx - 1], rtol=1e-12, atol=1e-14)
        step = r[:, 1] if nchk > 1 else np.zeros(r.shape[0])
        assert np.allclose(res_h[o][:, 1], step, rtol=1e-12, atol=1e-14)
    assert np.allclose(err_h[:sx + 1], err_r[:sx + 1], rtol=1e-12, atol=1e-300)


def test_sampling_error_is_standard_error_of_mean():
    """sigma of xsim.m:94-98 = sem over the e2*e3 sub-ensemble values."""
    e2, e3 = 6, 2
    p = _project(e2, e3, 1, 0, observe=["abs2_ave"])
    q = preferences(p)
    info = _info(q, [1])
    rng = np.random.default_rng(3)
    npts = q["points"][0]
    vals = rng.standard_normal((e3, e2, npts)) * 0.3 + 2.0
    data = None
    for npar in range(e3):
        av_all = np.zeros((e2, 2, npts), complex)
        av_all[:, 0] = vals[npar]
        av_all[:, 1] = vals[npar] + 0.01
        dp = ensemble_moments(q, info, [lambda av, p: av[0]], av_all)
        data = dp if data is None else [a + b for a, b in zip(data, dp)]
    _, res = statistics(q, data)
    flat = vals.reshape(e2 * e3, npts)
    assert np.allclose(res[0][:, 0], flat.mean(axis=0), rtol=1e-13)
    assert np.allclose(res[0][:, 2], stats.sem(flat, axis=0), rtol=1e-10)
    assert np.allclose(res[0][:, 1], 0.01, rtol=1e-9)
    # a sigma without the 1/(n-1) (population sd / sqrt(n)) would be smaller by sqrt((n-1)/n)
    assert not np.allclose(res[0][:, 2], np.std(flat, axis=0) / np.sqrt(e2 * e3), rtol=1e-3)


def test_richardson_extrapolation_order2():
    """D_coarse = X + c h^2, D_fine = X + c (h/2)^2, order 2: the extrapolated mean
    is X exactly and the step error |D_fine - X| = c (h/2)^2."""
    e2 = 3
    p = _project(e2, 1, 1, 2, observe=["abs2_ave"])
    q = preferences(p)
    info = _info(q, [1])
    npts = q["points"][0]
    t = np.linspace(0, 1, npts)
    exact = np.cos(t)
    c, h = 0.7 + 0.2 * t, 0.1
    av_all = np.zeros((e2, 2, npts), complex)
    av_all[:, 0] = exact + c * (h / 2) ** 2
    av_all[:, 1] = exact + c * h ** 2
    data = ensemble_moments(q, info, [lambda av, p: av[0]], av_all)
    _, res = statistics(q, data)
    assert np.allclose(res[0][:, 0], exact, rtol=0, atol=1e-14)
    assert np.allclose(res[0][:, 1], c * (h / 2) ** 2, rtol=1e-10)
    # order 1 would give a different mean (the test discriminates)
    q1 = dict(q, order=1)
    _, res1 = statistics(q1, ensemble_moments(q1, info, [lambda av, p: av[0]], av_all))
    assert np.abs(res1[0][:, 0] - exact).max() > 1e-4


def test_split_averages_time_fastest():
    """A flat average vector in MATLAB column-major order [lines, points, sites]
    (time fastest) splits into av[o] = [1, points, sites]."""
    q = preferences(_project(1, 1, 0, 0, nx=4))
    info = _info(q, [4, 1])
    npts = q["points"][0]
    a0 = np.arange(npts * 4).reshape(1, npts, 4) + 0j
    a1 = 100 + np.arange(npts).reshape(1, npts, 1) + 0j
    flat = np.concatenate([a0.reshape(-1, order="F"), a1.reshape(-1, order="F")])
    av = split_averages(q, info, flat)
    assert np.array_equal(av[0], a0) and np.array_equal(av[1], a1)
    assert flat[1] == a0[0, 1, 0]           # time fastest


def test_default_noise_and_random_counts_match_the_oracle():
    """xpreferences.m:65-76: noises default to the field count only when no
    noise cell is named (an explicit 0 k-noises names one), inrandoms to
    noises only when no random cell is named; the host and the oracle agree."""
    from oracle import xspde as X
    base = dict(dimensions=2, points=[3, 32], gpu=dict(deriv="zero", observe=["a"]))
    for extra in ({}, dict(knoises=0), dict(knoises=2), dict(urandoms=1), dict(krandoms=0), dict(noises=2),
                  dict(fields=2), dict(noises=2, inrandoms=1)):
        p = dict(base, **extra)
        q = preferences(p)
        o = X.xpreferences(dict(p, gpu=None))
        assert (q["noises"], q["inrandoms"]) == (o.noises, o.inrandoms), extra


@pytest.mark.parametrize("e2,e3,checks,order", [(3, 2, 1, 0), (4, 1, 1, 2), (5, 1, 0, 0)])
def test_host_comparisons_match_oracle(e2, e3, checks, order):
    """p.compare (xsim.m:110-143): the difference error, the chi-squared sum
    and the normalisation by the comparison's maximum, for a per-site output
    and a lattice-mean output, equal the oracle's xsim_stats on the same data."""
    cmp = [lambda p: 1.0 + 0.3 * np.cos(p.t) * np.exp(-p.x ** 2), lambda p: 2.0 + p.t]
    p = dict(_project(e2, e3, checks, order), compare=cmp)
    q = preferences(p)
    nx = q["points"][1]
    ns_list = [nx, 1]
    info = _info(q, ns_list)
    outs = [lambda av, p: av[0], OUTPUTS[1]]
    nchk = 1 + checks
    flat, arrays = _synthetic_averages(q, ns_list, e2, nchk, seed=7 + e2)
    data = None
    for npar in range(e3):
        dp = ensemble_moments(q, info, outs, flat[npar])
        data = dp if data is None else [a + b for a, b in zip(data, dp)]
    zero = split_averages(q, info, np.zeros(info.avlen, dtype=complex))
    shapes = [np.asarray(out(zero, q)).shape for out in outs]
    err_h, _ = statistics(q, data, shapes)
    po = X.xpreferences(dict(p, observe=[lambda a, p: a[0:1], lambda a, p: np.mean(a[0:1], axis=1)],
                             output=[None, OUTPUTS[1]]))
    dref = None
    for npar in range(1, e3 + 1):
        dp = X.xensemble(npar, po, path_fn=lambda p_, j, nc, n_: arrays[n_ - 1][(j, nc)])
        dref = dp if dref is None else [a + b for a, b in zip(dref, dp)]
    po._avshape = [(1, q["points"][0], nx), (1, q["points"][0])]
    err_r, _ = X.xsim_stats(po, dref)
    sx = max(3, 1 + nchk)
    print("host %s oracle %s" % (np.array2string(err_h[:sx + 2], precision=6), np.array2string(err_r[:sx + 2], precision=6)))
    assert err_h[sx] > 0                                   # the difference error is there
    assert np.allclose(err_h[:sx + 2], err_r[:sx + 2], rtol=1e-12, atol=1e-300)
