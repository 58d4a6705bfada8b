"""Projects with several field components (p.fields > 1, xpreferences.m:44-50)
on the GPU.

The built-in drift forms act on each component by itself, so a project with F
components runs as F one-component handles launched together, each with its
own row of p.boundaries{d} / p.boundval{d} (spdpreferences.m:17,26-66), its
(lower, upper) pair of a per-field p.boundfun, its share of the noise
components and its own counter stream (instance k).  Checked against:
  * the oracle's genuine multi-field problems: Boundaries2.m dataset (I)
    (five heat-equation fields with their own boundary types; closed forms to
    rounding, as in the oracle) and Peregrine.m as one 4-field problem
    (Batchtest #42, 4.588574e-04 at 0.1 %);
  * the oracle path by path, per component with the same instance stream, for
    stochastic lattice and SDE projects (FP64 1e-10, FP32 5e-5).
"""
import numpy as np
import pytest

from oracle import xspde as X
from xspde_matlab_amd.xsim import FieldsProblem, Problem, field_problems, preferences, xsim

from _problems import oracle_problem

pytestmark = pytest.mark.gpu


def _cmp(gpu, ref, rtol, what):
    gpu, ref = np.asarray(gpu), np.asarray(ref)
    err = np.abs(gpu - ref).max() / max(np.abs(ref).max(), 1e-300)
    print("%s: max |gpu-oracle|/max|oracle| = %.3e (tol %.1e)" % (what, err, rtol), flush=True)
    assert err <= rtol, what


def heat_fields(x):
    return np.stack([4 * np.sin(x) + np.sin(2 * x), 5 + 4 * np.cos(x) + np.cos(2 * x),
                     4 * np.sin(x / 2) + np.sin(3 * x / 2), 4 * np.cos(x / 2) + np.cos(3 * x / 2),
                     2 + np.cos(2 * x / 1.02) + np.sin(4 * x / 1.02)])


HEAT_EXACT = [
    lambda p: 4 * np.sin(p.x) * np.exp(-p.t) + np.sin(2 * p.x) * np.exp(-4 * p.t),
    lambda p: 5 + 4 * np.cos(p.x) * np.exp(-p.t) + np.cos(2 * p.x) * np.exp(-4 * p.t),
    lambda p: 4 * np.sin(p.x / 2) * np.exp(-p.t / 4) + np.sin(3 * p.x / 2) * np.exp(-9 * p.t / 4),
    lambda p: 4 * np.cos(p.x / 2) * np.exp(-p.t / 4) + np.cos(3 * p.x / 2) * np.exp(-9 * p.t / 4),
    lambda p: 2 + np.cos(2 * p.x / 1.02) * np.exp(-4 * p.t / 1.02 ** 2)
    + np.sin(4 * p.x / 1.02) * np.exp(-16 * p.t / 1.02 ** 2)]

BOUNDARIES2 = [[1, 1], [-1, -1], [1, -1], [-1, 1], [0, 0]]


def test_boundaries2_five_fields_spectral():
    """Examples/SPDE_2D/Boundaries2.m dataset (I): five fields of the heat
    equation, DD, NN, DN, ND and periodic, MP, 51 x 51 (padded to the 64-site
    FP64 dense-transform kernel; the odd periodic field on a padded DFT)."""
    p = dict(name="Boundaries2", dimensions=2, points=[51, 51], method="MP", order=0, fields=5,
             ranges=[4, np.pi], origins=[0, 0], boundaries=[BOUNDARIES2], ensembles=[1, 1, 1],
             gpu=dict(deriv="zero", linear=dict(l2=1.0), initial="const", observe=["a"] * 5,
                      ofield=[1, 2, 3, 4, 5], precision="fp64"))
    prob = FieldsProblem(p)
    q = X.xpreferences(dict(dimensions=2, points=[51, 51], method="MP", order=0, fields=5, ranges=[4, np.pi],
                            origins=[0, 0], boundaries=[BOUNDARIES2], linear=lambda p: p.Dx ** 2,
                            initial=lambda v, p: heat_fields(p.x[0]), compare=HEAT_EXACT,
                            observe=[lambda a, p, _k=k: a[_k][None] for k in range(5)]))
    a0 = heat_fields(q.xc[1])[:, None, :]                             # [F, E, sites]
    gpu_av = {}
    for nc in (1, 2):
        gpu_av[nc], a_g = prob.xpath(nc, npar=1, j=1, a0=a0, want_final=True)
        a_r, av_r = X.run_path(q, 1, nc)
        _cmp(a_g[:, 0, :], a_r[:, :, 0], 1e-10, "Boundaries2 final fields nc=%d" % nc)
        for o in range(5):
            _cmp(gpu_av[nc][o], av_r[o], 1e-10, "Boundaries2 field %d nc=%d" % (o + 1, nc))
    prob.close()
    q._avshape = X._probe_shapes(q)
    data = X.xensemble(1, q, path_fn=lambda p_, j, nc, npar: gpu_av[nc])
    err = X.xsim_stats(q, data)[0]
    print("Boundaries2 (I) on the GPU: error %.3e" % err[0], flush=True)
    assert err[0] < 1e-12


def _peregrine(x, t):
    return np.exp(1j * t) * (4 * (1 + 2j * t) / (1 + 4 * (t ** 2 + x ** 2)) - 1)


def _peregrine_dx(x, t):
    return -8 * x * np.exp(1j * t) * (4 * (1 + 2j * t) / (1 + 4 * (t ** 2 + x ** 2)) ** 2)


def _peregrine_ends(a, c, d, p):
    """One (lower, upper) pair per field, DD, NN, DN, ND (Peregrine.m:45-50)."""
    v, dv = _peregrine(p.origins[1], p.t), _peregrine_dx(p.origins[1], p.t)
    return [(v, v), (dv, -dv), (v, -dv), (dv, v)]


def test_peregrine_as_one_four_field_project():
    """Examples/SPDE_2D/Peregrine.m as the reference writes it (Batchtest #42:
    4.588574e-04, tolerance 0.1 %): one project of four fields, one boundary
    row per field, a boundary function returning one pair per field; 161
    sites on the 256-site FP64 dense-transform kernel, MP, order 2."""
    p = dict(name="Peregrine", dimensions=2, noises=0, fields=4, ranges=[10.0, 10.0], origins=[-5.0, -5.0],
             points=[51, 161], method="MP", order=2, steps=20, checks=1, ensembles=[1, 1, 1],
             boundaries=[[[1, 1], [-1, -1], [1, -1], [-1, 1]]], boundfun=_peregrine_ends,
             gpu=dict(deriv="gp", g=1j, s=0.0, linear=dict(l2=0.5j), initial="const", c0=0.0,
                      observe=["a"] * 4, ofield=[1, 2, 3, 4], precision="fp64"))
    prob = FieldsProblem(p)
    q = X.xpreferences(dict(dimensions=2, noises=0, fields=4, ranges=[10, 10], origins=[-5, -5], points=[51, 161],
                            method="MP", order=2, steps=20, boundaries=p["boundaries"], boundfun=_peregrine_ends,
                            initial=lambda v, p: _peregrine(p.x, p.origins[0]) + np.zeros((4, 1, 1)),
                            deriv=lambda a, w, p: 1j * a * (np.conj(a) * a), linear=lambda p: 0.5j * p.Dx ** 2,
                            observe=[lambda a, p, _k=k: a[_k][None] for k in range(4)],
                            output=[lambda o, p, _k=k: np.abs(o[_k]) ** 2 for k in range(4)],
                            compare=[lambda p: np.abs(_peregrine(p.x, p.t)) ** 2] * 4))
    a0 = np.broadcast_to(_peregrine(q.xc[1], -5.0), (4, 1, 161))
    gpu_av = {nc: prob.xpath(nc, npar=1, j=1, a0=a0) for nc in (1, 2)}
    prob.close()
    q._avshape = X._probe_shapes(q)
    data = X.xensemble(1, q, path_fn=lambda p_, j, nc, npar: gpu_av[nc])
    err = X.xsim_stats(q, data)[0]
    print("Peregrine (4 fields) on the GPU: error %.7e (Batchtest 4.588574e-04)" % err[0], flush=True)
    assert 4.588574e-04 / 1.001 < err[0] < 4.588574e-04 * 1.001


def _oracle_fields(p, j, nc):
    """The oracle path of every component: its one-field project with instance k's streams."""
    subs, where = field_problems(p, preferences(p))
    res = []
    for k, sk in enumerate(subs):
        q = X.xpreferences(oracle_problem(sk))
        res.append(X.run_path(q, j, nc, 1, inst=k))
    return res, where


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 5e-5)])
def test_stochastic_lattice_fields_match_oracle(prec, tol):
    """Two stochastic CGL components on a periodic 64-site lattice with their own
    drift parameters (p.gpu.fields overrides) and two noise components each:
    every component's final fields and observables equal the oracle's path with
    that component's instance stream; the assembled observables keep the
    project's order."""
    p = dict(name="TwoFields", dimensions=2, fields=2, noises=4, points=[6, 64], ranges=[0.5, 20.0], steps=[4, 1],
             checks=1, ensembles=[8, 2, 1], method="RK4", seed=5,
             gpu=dict(deriv="gp", g=-1 - 0.3j, u=0.1, s=0.05, linear=dict(l0=0.2, l2=0.5j), initial="sech",
                      amp=1.0, width=2.0, eps=0.1, observe=["abs2_ave", "a", "abs2"], ofield=[1, 2, 1],
                      fields=[{}, dict(g=-0.5 + 0.2j, linear=dict(l0=-0.1, l2=0.8j), amp=0.7)], precision=prec))
    prob = FieldsProblem(p)
    for j, nc in ((1, 1), (2, 2)):
        av_g, a_g = prob.xpath(nc, npar=1, j=j, want_final=True)
        ref, where = _oracle_fields(p, j, nc)
        for k in range(2):
            a_r = ref[k][0]
            _cmp(a_g[k], np.moveaxis(a_r[0], -1, 0), tol, "%s field %d final j=%d nc=%d" % (prec, k + 1, j, nc))
        for o, (k, i) in enumerate(where):
            _cmp(av_g[o], ref[k][1][i], tol, "%s observable %d (field %d) j=%d nc=%d" % (prec, o + 1, k + 1, j, nc))
    prob.close()


def test_sde_fields_and_xsim():
    """Two Kubo components (dimensions 1) with their own c, one noise each:
    paths against the oracle per component, then xsim's statistics run over the
    assembled outputs (one data block per observable, in the project's order)."""
    p = dict(name="Kubo2", dimensions=1, fields=2, noises=2, points=[21], ranges=[2.0], steps=[5], checks=1,
             ensembles=[128, 16, 1], method="MP", seed=11,
             gpu=dict(deriv="kubo", c=1.0, initial="const", c0=1.0, observe=["a", "abs2", "a"], ofield=[1, 1, 2],
                      fields=[{}, dict(c=0.5)], precision="fp64"))
    prob = FieldsProblem(p)
    av_g = prob.xpath(2, npar=1, j=3)
    ref, where = _oracle_fields(p, 3, 2)
    for o, (k, i) in enumerate(where):
        _cmp(av_g[o], ref[k][1][i], 1e-10, "Kubo2 observable %d (field %d)" % (o + 1, k + 1))
    prob.close()
    err, data, q, _ = xsim(p)
    assert len(data) == 3 and all(np.all(np.isfinite(d)) for d in data)
    # <a(t)> = exp(-c^2 t / 2) for Kubo (the c = 1 and c = 0.5 components), within sampling errors
    t = np.linspace(0, 2.0, 21)
    for o, c in ((0, 1.0), (2, 0.5)):
        z = (data[o][:, 0] - np.exp(-c * c * t / 2)) / np.maximum(data[o][:, 2], 1e-12)
        print("Kubo2 field c=%.1f: max |z| %.2f" % (c, np.abs(z[1:]).max()), flush=True)
        assert np.abs(z[1:]).max() < 5


# ---------------------------------------------------------------- coupled SDE components
# The coupled-component kernel (sde_fields) was written after GPU access closed
# for this build's last session: these parity tests have not run on an MI355X.
# They stay skipped until they do; the oracle side is pinned on the CPU
# (tests/test_fields_cpu.py: the SHO closed form, descriptor planning, refusals).
COUPLED_NOT_RUN = pytest.mark.skip(reason="sde_fields has not run on an MI355X yet (no GPU access after it was "
                                          "written); oracle side pinned in test_fields_cpu.py")


def sho(prec, method="MP"):
    """Examples/SDE_1D/SHO.m:9-20: da = (-a2, a1) dt from (1, 0), 20 steps per
    point, order 2, ensembles 3 (MP by the xpreferences default)."""
    return dict(name="SHO", dimensions=1, fields=2, noises=0, points=[51], ranges=[10.0], steps=[20], checks=1,
                order=2, ensembles=[3, 1, 1], method=method,
                gpu=dict(deriv="matrix", M=[[0, -1], [1, 0]], initial="const", c0=[1.0, 0.0],
                         observe=["a", "a", "abs2"], ofield=[1, 2, 1], precision=prec))


@COUPLED_NOT_RUN
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
@pytest.mark.parametrize("method", ["MP", "RK4"])
def test_coupled_sho_matches_oracle_and_cosine(prec, tol, method):
    """Coupled components on the GPU (one descriptor, the matrix drift form)
    against the oracle path, fine and coarse, and the extrapolated ensemble mean
    of a_1 against the closed form cos(t) (the compare of SHO.m:16)."""
    p = sho(prec, method)
    prob = Problem(p)
    assert prob.info.nsites == 2 and prob.info.kernel_kind == 3
    q = X.xpreferences(oracle_problem(p))
    gpu_av = {}
    for nc in (1, 2):
        av_g, a_g = prob.xpath(nc, npar=1, j=1, want_final=True)
        a_r, av_r = X.run_path(q, 1, nc)
        _cmp(np.asarray(a_g).reshape(3, 2), a_r.T, tol, "SHO %s %s final fields nc=%d" % (prec, method, nc))
        for o in range(3):
            _cmp(av_g[o], av_r[o], tol, "SHO %s %s observable %d nc=%d" % (prec, method, o + 1, nc))
        gpu_av[nc] = av_g
    prob.close()
    q.compare = [lambda p: np.cos(p.t), lambda p: np.sin(p.t), lambda p: np.cos(p.t) ** 2]
    q._avshape = X._probe_shapes(q)
    data = X.xensemble(1, q, path_fn=lambda p_, j, nc, npar: gpu_av[nc])
    err = X.xsim_stats(q, data)[0]
    print("SHO %s %s on the GPU: error vector %s" % (prec, method, np.array2string(err[:5], precision=3)), flush=True)
    # MP's extrapolated error vs cos(t) is 5.6e-6 (oracle), RK4's 9e-11
    assert err[0] < (1e-4 if method == "MP" else (1e-8 if prec == "fp64" else 2e-6))


@COUPLED_NOT_RUN
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 5e-5)])
@pytest.mark.parametrize("method", ["RK4", "MP"])
def test_coupled_stochastic_components_match_oracle(prec, tol, method):
    """Three coupled components with a complex coupling matrix, a nonlinearity
    per component, propagators per component and two noise components mixed by
    S: every trajectory's final components and every observable against the
    oracle path (same counter stream), fine and coarse, two sub-ensembles."""
    M = [[-0.2 + 0.5j, 0.3, 0.0], [-0.3, -0.1j, 0.2 + 0.1j], [0.1j, -0.2, -0.3]]
    p = dict(name="Coupled3", dimensions=1, fields=3, noises=2, points=[6], ranges=[1.5], steps=[3], checks=1,
             ensembles=[300, 2, 1], method=method, seed=9,
             gpu=dict(deriv="matrix", M=M, g=[-0.2j, 0.1, -0.05 + 0.1j], S=[[0.3, 0.0], [0.1, 0.2], [0.0, -0.25]],
                      linear=dict(l0=[-0.1 + 0.4j, 0.2j, -0.05]), initial="const", c0=[1.0, 0.5j, -0.3 + 0.2j],
                      observe=["a", "abs2", "a_ave", "abs4_ave"], ofield=[3, 1, 2, 3], precision=prec))
    prob = Problem(p)
    q = X.xpreferences(oracle_problem(p))
    for j, nc in ((1, 1), (2, 2)):
        av_g, a_g = prob.xpath(nc, npar=1, j=j, want_final=True)
        a_r, av_r = X.run_path(q, j, nc)
        _cmp(np.asarray(a_g).reshape(300, 3), a_r.T, tol, "Coupled3 %s %s final j=%d nc=%d" % (prec, method, j, nc))
        for o in range(4):
            _cmp(av_g[o], av_r[o], tol, "Coupled3 %s %s observable %d j=%d nc=%d" % (prec, method, o + 1, j, nc))
    prob.close()
    err, data, _, _ = xsim(p)
    assert len(data) == 4 and all(np.all(np.isfinite(d)) for d in data)
