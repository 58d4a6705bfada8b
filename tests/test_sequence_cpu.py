"""xsim of a sequence of problems on the host (xsim.m:33-38, xensemble.m:25-108,
xpreferences.m:11-13,29-32,197) without a device: the path of every problem
is the oracle's (a stand-in handle), so the product's orchestration - problem
order per (j, nc), default and custom p.transfer, the chained time origins,
the per-problem random streams, the moments per problem and the one error
vector over all outputs - must equal the oracle's own sequence xsim."""
import importlib

import numpy as np

from oracle import forms
from oracle import xspde as X

XS = importlib.import_module("xspde_matlab_amd.xsim")


class _OracleProblem:
    """Problem's interface with the oracle's path in place of the GPU handle."""

    def __init__(self, p):
        self.q = XS.preferences(p)
        self.qo = X.xpreferences(forms.apply_forms(p))
        self.seq = int(self.q["gpu"].get("seq", 0))
        n = len(self.q["gpu"].get("observe", ["a"]))
        self.outputs = [lambda av, p, _n=k: av[_n] for k in range(n)]

    def xpath(self, nc, npar=1, j=1, a0=None, want_final=False):
        a0o = None if a0 is None else XS._to_cells(self.q, a0)
        a, av = X.run_path(self.qo, j, nc, npar, seq=self.seq, a0=a0o)
        return av, XS._from_cells(self.q, a).reshape(-1)

    def xensemble(self, npar, want_raw=False):
        q = self.q
        avs = {(j, nc): self.xpath(nc + 1, npar, j + 1)[0]
               for j in range(q["ensembles"][1]) for nc in range(1 + q["checks"][0])}
        return XS.ensemble_moments(q, None, self.outputs, None, split=lambda j, nc: avs[(j, nc)])

    def zero_av(self):
        av, _ = self.xpath(1)
        return [np.zeros_like(x) for x in av]

    def close(self):
        pass


def _nls(g, **kw):
    p = dict(dimensions=2, points=[7, 32], ranges=[0.6, 12.0], steps=3, checks=1, ensembles=[4, 3, 2], noises=2,
             inrandoms=2, method="RK4", seed=4,
             gpu=dict(deriv="gp", g=g, u=-0.1, s=0.05, linear=dict(l0=-0.1, l2=0.5j), initial="sech", amp=1.0,
                      width=1.5, eps=0.1, observe=["abs2", "a_ave", "abs2_ave"], precision="fp64"))
    p.update(kw)
    return p


def test_sequence_matches_the_oracle_sequence(monkeypatch):
    monkeypatch.setattr(XS, "Problem", _OracleProblem)
    half = lambda b, *rest: 0.5 * b                    # p.transfer(b, v, p): custom transfer
    seq = [_nls(1j), _nls(2j, transfer=half), _nls(-1j, method="MP")]
    err, data, qs = XS.xsim_sequence(seq)
    err_o, data_o, qo = X.xsim([forms.apply_forms(p) for p in seq])
    # chained time origins (xpreferences.m:197) and the first problem's checks
    assert [q["origins"][0] for q in qs] == [0.0, 0.6, 1.2] == [float(q.origins[0]) for q in qo]
    sx = 3
    assert np.allclose(err[:sx + 1], err_o[:sx + 1], rtol=1e-12, atol=0), (err[:sx + 1], err_o[:sx + 1])
    for s in range(3):
        for o in range(3):
            d, r = data[s][o], data_o[s][o].reshape(data[s][o].shape[0], -1)
            assert np.allclose(d[:, 0], r[:, 0], rtol=1e-12, atol=1e-14)          # means
            assert np.allclose(d[:, 2], r[:, sx - 1], rtol=1e-10, atol=1e-14)     # sampling errors


def test_sequence_default_transfer_continues_the_field(monkeypatch):
    """Without p.transfer the second problem starts from the first one's final
    fields (xpreferences.m:94): a sequence of two identical noiseless linear
    problems equals one problem over both time ranges at the shared points
    (one step per point: with more, the final fields lie steps - 1 steps past
    the last stored point, xpath.m:29-86)."""
    monkeypatch.setattr(XS, "Problem", _OracleProblem)
    p = dict(dimensions=2, points=[5, 16], ranges=[0.4, 8.0], steps=1, checks=0, ensembles=[1, 1, 1], noises=0,
             gpu=dict(deriv="zero", linear=dict(l0=-0.3, l2=0.2j), initial="gauss", amp=1.0, width=1.0,
                      observe=["a"], precision="fp64"))
    _, data, _ = XS.xsim_sequence([p, dict(p)])
    whole = dict(p, points=[9, 16], ranges=[0.8, 8.0])
    qo = X.xpreferences(forms.apply_forms(whole))
    a, av = X.run_path(qo, 1, 1)
    second = data[1][0][:, 0].reshape(5, 16)                   # [points, x] of the second problem
    assert np.allclose(second, np.real(av[0][0, 4:, :]), rtol=1e-12, atol=1e-14)


def _soliton():
    """Examples/SPDE_2D/Soliton.m (Batchtest #4) in built-in forms, with its compare."""
    sech = lambda x: 1 / np.cosh(x)
    return dict(name="Soliton", dimensions=2, checks=0, order=2, noises=0, inrandoms=0,
                compare=[lambda p: sech(p.x)],
                gpu=dict(deriv="gp", g=1j, s=0.0, linear=dict(l0=-0.5j, l2=0.5j), initial="sech", amp=1.0,
                         width=1.0, eps=0.0, observe=["a"], precision="fp64"))


def test_host_xsim_reproduces_soliton_known_answer(monkeypatch):
    """The host's own xsim statistics with the comparison (xsim.m:110-188) on
    the oracle's paths give Batchtest #4's error, 5.047219e-03 within 0.1 %."""
    monkeypatch.setattr(XS, "Problem", _OracleProblem)
    err = XS.xsim(_soliton())[0]
    print("Soliton via the host xsim: %.7e" % err[0])
    assert 5.047219e-03 / 1.001 < err[0] < 5.047219e-03 * 1.001
