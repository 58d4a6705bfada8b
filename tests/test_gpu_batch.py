"""GPU tests of batched uneven instances (one graph launch, one persistent
kernel per lattice family): every instance's averages equal those of the same
project run on its own handle."""
import numpy as np
import pytest

from xspde_matlab_amd import Problem
from xspde_matlab_amd.configs import batch_instances
from xspde_matlab_amd.xsim import BatchProblem

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
