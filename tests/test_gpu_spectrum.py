"""GPU time-domain spectra (xpath.m:45-68,90-106) through the C ABI.

The SDE kernel keeps every step's field; a second kernel takes the points(1)
frequencies of fsc * ifft over time and the flagged observables of them.
Checked against the closed-form geometric series of a linear SDE (exact
interaction-picture steps) and against the oracle on a seeded
Ornstein-Uhlenbeck ensemble (Examples/SDE_1D/Equilibrium.m:12-28), fine and
coarse runs, FP64 within 1e-10 and FP32 within 2e-5 relative."""

import numpy as np
import pytest

from oracle import xspde as X
from xspde_matlab_amd import Problem, xsim

from _problems import oracle_problem

pytestmark = pytest.mark.gpu


def _cmp(gpu, ref, rtol, what):
    gpu, ref = np.asarray(gpu), np.asarray(ref)
    err = np.abs(gpu - ref).max() / max(np.abs(ref).max(), 1e-300)
    print("%s: max |gpu-ref|/max|ref| = %.3e (tol %.1e)" % (what, err, rtol), flush=True)
    assert err <= rtol, what


def linear_sde(prec, L, pts=21, steps=3, checks=0):
    return dict(name="LinearSpectrum", dimensions=1, points=[pts], ranges=[3.0], steps=[steps], checks=checks,
                noises=0, method="RK4", ensembles=[4, 1, 1], transforms=[0, 1],
                gpu=dict(deriv="zero", linear=dict(l0=L), initial="const", c0=1.0, observe=["a", "a"],
                         precision=prec))


def equilibrium(prec, E=256, e2=2, checks=1):
    """Examples/SDE_1D/Equilibrium.m:12-28 in built-in forms: da = (-a + w1 + i w2) dt,
    a(0) = (v1 + i v2)/sqrt(2); observables |a|^2 and |a(w)|^2."""
    return dict(name="Equilibrium", dimensions=1, points=[200], steps=[2], ranges=[25.0], seed=241, noises=2,
                inrandoms=2, checks=checks, ensembles=[E, e2, 1], method="MP", transforms=[0, 1],
                gpu=dict(deriv="gp", g=0.0, u=-1.0, s=1.0, initial="sech", amp=0.0, width=1.0, eps=1.0,
                         observe=["abs2", "abs2"], precision=prec))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 2e-5)])
@pytest.mark.parametrize("pts,steps", [(21, 3), (20, 2)])
def test_linear_sde_spectrum_closed_form(prec, tol, pts, steps):
    L = -0.3 + 1.1j
    p = linear_sde(prec, L, pts, steps)
    prob = Problem(p)
    av = prob.xpath(1, npar=1, j=1)
    prob.close()
    q = X.xpreferences(oracle_problem(p))
    M = pts * steps
    z = np.exp(L * q.dt / steps)
    qq = z * np.exp(2j * np.pi * q.plotpts / M)
    expect = q.fsc / M * (1 - qq ** M) / (1 - qq)
    _cmp(av[1][0, :, 0], expect, tol, "spectrum %s %dx%d" % (prec, pts, steps))
    _cmp(av[0][0, :, 0], z ** (steps * np.arange(pts)), tol, "time samples %s" % prec)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 2e-5)])
def test_ou_spectrum_matches_oracle(prec, tol):
    p = equilibrium(prec, E=96)
    prob = Problem(p)
    for j in (1, 2):
        for nc in (1, 2):
            av = prob.xpath(nc, npar=1, j=j)
            q = X.xpreferences(oracle_problem(p))
            _, av_r = X.run_path(q, j, nc)
            for o in range(2):
                _cmp(av[o], av_r[o], tol, "OU %s j=%d nc=%d av[%d]" % (prec, j, nc, o))
    prob.close()


def test_ou_spectrum_ensemble_and_statistics():
    """xsim on the GPU (all paths in one launch, then the spectrum kernel over
    every path) against the oracle's xsim, and the mean spectrum against
    Equilibrium.m's compare function ranges / (pi (1 + w^2))."""
    p = equilibrium("fp64", E=128, e2=16)
    err_g, data_g, q, _ = xsim(p)
    err_r, outd, qo = X.xsim(oracle_problem(p))
    sx = max(3, 1 + qo.nchk)
    for o in range(2):
        ref = outd[o].reshape(-1, outd[o].shape[-1])
        _cmp(data_g[o][:, 0], ref[:, 0], 1e-10, "mean o=%d" % o)
        _cmp(data_g[o][:, 2], ref[:, sx - 1], 1e-8, "sampling error o=%d" % o)
    w = qo.w
    S, sig = data_g[1][:, 0], data_g[1][:, 2]
    compare = qo.ranges[0] / (np.pi * (1 + w ** 2))
    sel = np.abs(w) < 3
    z = (S[sel] - compare[sel]) / sig[sel]
    print("OU spectrum: mean z %.3f, rms z %.3f over %d frequencies" % (z.mean(), np.sqrt((z ** 2).mean()), sel.sum()))
    # sigma from 16 sub-ensembles: z is t-distributed (15 degrees of freedom)
    assert abs(z.mean()) < 0.6 and np.sqrt((z ** 2).mean()) < 2.5
