"""GPU tests of batched uneven instances (one graph launch, one persistent
kernel per lattice family): every instance's averages equal those of the same
project run on its own handle, and, for one instance of every lattice family of
the config-4 batch (1D 32..512, 2D 32^2, 64^2, 128^2), those of the FP64 oracle's
xensemble (Simulator/xensemble.m:25-80) on the same counter streams, untuned
and autotuned."""
import numpy as np
import pytest

from oracle import xspde as X
from xspde_matlab_amd import Problem
from xspde_matlab_amd.configs import BATCH_1D, BATCH_2D, batch_instances
from xspde_matlab_amd.xsim import BatchProblem

from _problems import oracle_problem

pytestmark = pytest.mark.gpu


def _small_batch():
    ps = batch_instances(count=12, seed=3, steps=20, points=3)
    for p in ps:                       # keep the test short: at most 512 trajectories each
        E = min(p["ensembles"][0] * p["ensembles"][1], 512)
        p["ensembles"] = [E // 2, 2, 1]
    return ps


def _single(p):
    prob = Problem(p)
    prob.handle.run_ensemble(1)
    prob.handle.synchronize()
    av = prob.handle.read_averages()
    prob.close()
    return av


@pytest.mark.parametrize("autotune,serial,tol", [(False, False, 1e-12), (True, False, 2e-5), (False, True, 1e-12)])
def test_batch_matches_single_runs(autotune, serial, tol):
    ps = _small_batch()
    sizes = sorted({tuple(p["points"][1:]) for p in ps})
    print("lattices", sizes)
    assert len(sizes) >= 3
    bp = BatchProblem(ps, autotune=autotune, serial=serial)
    print("plan", bp.plan())
    for rep in range(2):               # replay the captured graph
        bp.run(1)
        ms = bp.synchronize()
        print("batch run %d: %.2f ms" % (rep, ms))
        for i, p in enumerate(ps):
            got = bp.read_averages(i)
            ref = _single(p)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err < tol, (i, p["points"], err)
    bp.close()


def test_batch_rejects_sde_instance():
    from xspde_matlab_amd.configs import kubo
    from xspde_matlab_amd._lib import XshError
    with pytest.raises(XshError):
        BatchProblem([kubo("fp32")] + _small_batch()[:1])


def _family_batch():
    """One instance per lattice family of the config-4 batch (its drift, damping
    and nonlinearity drawn by seed), small ensembles and few steps."""
    src = batch_instances(count=64, seed=0, steps=500)
    out = []
    for shape in [(n,) for n in BATCH_1D] + [(n, n) for n in BATCH_2D]:
        p = next((q for q in src if tuple(q["points"][1:]) == shape), None)
        if p is None:
            p = dict(src[0], dimensions=len(shape) + 1, points=[src[0]["points"][0]] + list(shape),
                     ranges=[src[0]["ranges"][0]] + [20.0] * len(shape), steps=[src[0]["steps"][0]] + [1] * len(shape))
        p = dict(p, gpu=dict(p["gpu"]))
        E = 3 if len(shape) == 2 else 6
        p["ensembles"] = [E, 2, 1]
        p["points"] = [3] + list(p["points"][1:])
        p["steps"] = [2] + list(p["steps"][1:])
        p["ranges"] = [0.005 * 2 * 2] + list(p["ranges"][1:])
        p["checks"] = 1
        out.append(p)
    return out


def _oracle_averages(p):
    """[e2, nchk, avlen] in the C ABI's order (lattice means: one value per point)."""
    q = X.xpreferences(oracle_problem(p))
    rows = []
    for j in range(1, q.ensembles[1] + 1):
        rows.append([np.concatenate([a.reshape(-1, order="F") for a in X.run_path(q, j, nc)[1]])
                     for nc in range(1, q.nchk + 1)])
    return np.array(rows)


@pytest.mark.parametrize("autotune", [False, True])
def test_batch_families_match_oracle(autotune):
    ps = _family_batch()
    assert {tuple(p["points"][1:]) for p in ps} == {(n,) for n in BATCH_1D} | {(n, n) for n in BATCH_2D}
    bp = BatchProblem(ps, autotune=autotune)
    print("plan", bp.plan())
    bp.run(1)
    bp.synchronize()
    for i, p in enumerate(ps):
        got = bp.read_averages(i)
        ref = _oracle_averages(p)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print("instance %d %s: max rel diff vs oracle %.2e" % (i, p["points"][1:], err), flush=True)
        assert err < 5e-5, (i, p["points"], err)
    bp.close()


def _kfilter(w, p):
    """A k-space noise filter neither real nor symmetric in k (xnoise.m:44)."""
    k = p.kx if p.dimensions == 2 else p.kx - 0.5 * p.ky
    return w * np.exp(-0.5 * (k - 0.3) ** 2) * (1 + 0.2j * k)


@pytest.mark.parametrize("autotune", [False, True])
def test_batch_with_kspace_noise(autotune):
    """k-space filtered noises (knoises, xnoise.m:39-50) in a batch: each
    instance keeps its x-space noise planes per work item of its own, whatever
    block of the persistent grid serves the item; instances with and without
    k-space noise, 1D and 2D, against their single runs and the oracle."""
    ps = _family_batch()
    pick = [p for p in ps if tuple(p["points"][1:]) in ((64,), (256,), (32, 32), (64, 64))]
    out = []
    for i, p in enumerate(pick):
        q = dict(p, gpu=dict(p["gpu"]))
        if i % 2 == 0:
            q.update(noises=0, knoises=2, nfilter=_kfilter)
        out.append(q)
    bp = BatchProblem(out, autotune=autotune)
    print("plan", bp.plan())
    for rep in range(2):
        bp.run(1)
        bp.synchronize()
        for i, p in enumerate(out):
            got = bp.read_averages(i)
            ref = _single(p)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print("rep %d instance %d %s knoises %d: vs single run %.2e" % (rep, i, p["points"][1:], p.get("knoises", 0), err),
                  flush=True)
            assert err < (2e-5 if autotune else 1e-12), (i, err)
            if rep == 0:
                oerr = np.abs(got - _oracle_averages(p)).max() / np.abs(ref).max()
                print("   vs oracle %.2e" % oerr, flush=True)
                assert oerr < 5e-5, (i, oerr)
    bp.close()
