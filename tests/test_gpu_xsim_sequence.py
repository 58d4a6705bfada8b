"""xsim of a sequence of problems on the GPU path (xsim.m:33-38,
xensemble.m:25-108): every problem's path runs on the device, each after the
first starting from p.transfer of the previous problem's final fields; the
error vector over all outputs and every problem's means and sampling errors
against the oracle's sequence xsim (same counter streams, problem s keyed by
s), FP64 within 1e-9 (means) and 1e-6 (error vector) relative."""
import numpy as np
import pytest

from oracle import forms
from oracle import xspde as X
from xspde_matlab_amd.xsim import xsim

pytestmark = pytest.mark.gpu


def _nls(g, **kw):
    p = dict(dimensions=2, points=[7, 32], ranges=[0.6, 12.0], steps=3, checks=1, ensembles=[4, 3, 2], noises=2,
             inrandoms=2, method="RK4", seed=4,
             gpu=dict(deriv="gp", g=g, u=-0.1, s=0.05, linear=dict(l0=-0.1, l2=0.5j), initial="sech", amp=1.0,
                      width=1.5, eps=0.1, observe=["abs2", "a_ave", "abs2_ave"], precision="fp64"))
    p.update(kw)
    return p


def test_sequence_on_the_gpu_matches_the_oracle():
    half = lambda b, *rest: 0.5 * b
    seq = [_nls(1j), _nls(2j, transfer=half), _nls(-1j, method="MP")]
    err, data, qs, raw = xsim(seq)
    err_o, data_o, _ = X.xsim([forms.apply_forms(p) for p in seq])
    print("sequence error vector gpu %s oracle %s" % (np.array2string(err[:4], precision=6),
                                                      np.array2string(err_o[:4], precision=6)), flush=True)
    assert raw is None and len(data) == 3
    assert np.allclose(err[:4], err_o[:4], rtol=1e-6, atol=0)
    for s in range(3):
        for o in range(3):
            d, r = data[s][o], data_o[s][o].reshape(data[s][o].shape[0], -1)
            assert np.allclose(d[:, 0], r[:, 0], rtol=1e-9, atol=1e-12), (s, o)
            assert np.allclose(d[:, 2], r[:, 2], rtol=1e-6, atol=1e-12), (s, o)


def test_host_xsim_on_the_gpu_reproduces_soliton():
    """Examples/SPDE_2D/Soliton.m (Batchtest #4, 5.047219e-03 at 0.1 %) through
    the host's own xsim end to end: the built-in NLS forms on the GPU (35
    periodic sites on the padded FP64 DFT kernel), the host's moments and its
    statistics with the comparison sech(x)."""
    sech = lambda x: 1 / np.cosh(x)
    p = dict(name="Soliton", dimensions=2, checks=0, order=2, noises=0, inrandoms=0,
             compare=[lambda p: sech(p.x)],
             gpu=dict(deriv="gp", g=1j, s=0.0, linear=dict(l0=-0.5j, l2=0.5j), initial="sech", amp=1.0,
                      width=1.0, eps=0.0, observe=["a"], precision="fp64"))
    err = xsim(p)[0]
    print("Soliton via the host xsim on the GPU: %.7e (Batchtest 5.047219e-03)" % err[0], flush=True)
    assert 5.047219e-03 / 1.001 < err[0] < 5.047219e-03 * 1.001
