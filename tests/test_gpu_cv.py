"""GPU tests of the learned control variate (truncated-Wigner configuration):
the MFMA network against the FP64 oracle (oracle/mlp.py with bf16 operand
rounding), the on-device Gaussian sampler against the oracle's stream, the
per-trajectory observables and noise modes against the oracle path, and the
variance reduction of the trained estimator."""
import math

import numpy as np
import pytest

from oracle import mlp as OM
from oracle import philox
from oracle import xspde as X
from xspde_matlab_amd import _lib, cv
from xspde_matlab_amd.configs import shrink, tw2d
from xspde_matlab_amd.xsim import Problem, lowest_modes

from _problems import oracle_problem

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def test_mlp_forward_and_gradient_match_oracle():
    rng = np.random.default_rng(5)
    th = OM.init(1).astype(np.float32)
    net = _lib.MLP(th, batch_max=512)
    x = rng.standard_normal((300, 64)).astype(np.float32)
    t = rng.standard_normal(300).astype(np.float32)
    y = net.forward(x)
    yr, _ = OM.forward(th.astype(np.float64), x.astype(np.float64), rb=True)
    e = _rel(y, yr)
    print("forward rel err %.2e" % e)
    assert e < 2e-3
    loss, g = net.loss_grad(x, t)
    lr, gr = OM.loss_and_grad(th.astype(np.float64), x.astype(np.float64), t.astype(np.float64), rb=True)
    assert abs(loss - lr) / lr < 1e-3
    o = 0
    for name, sh in zip(["W1", "b1", "W2", "b2", "w3", "b3"], OM.sizes()):
        k = int(np.prod(sh))
        e = _rel(g[o:o + k], gr[o:o + k])
        print("grad %s rel err %.2e" % (name, e))
        assert e < 5e-3, name
        o += k
    net.close()


def test_training_reduces_loss():
    """Adam on the device fits a nonlinear function of 64 normal features."""
    rng = np.random.default_rng(2)
    z = rng.standard_normal((16384, 64)).astype(np.float32)
    y = (z[:, 0] + 0.5 * z[:, 1] ** 2 - 0.3 * z[:, 2] * z[:, 3]).astype(np.float32)
    m = cv.CVModel(seed=3, batch_max=4096)
    info = m.fit(z, y, epochs=15, batch=256, lr=2e-3, linear=False, val_frac=0.0)
    losses = info["losses"]
    print("loss first %.3f last %.4f" % (losses[0], losses[-1]))
    assert losses[-1] < 0.1 * losses[0]
    zt = rng.standard_normal((4096, 64)).astype(np.float32)
    yt = zt[:, 0] + 0.5 * zt[:, 1] ** 2 - 0.3 * zt[:, 2] * zt[:, 3]
    resid = np.var(yt - m.predict(zt)) / np.var(yt)
    print("held-out residual variance fraction %.4f" % resid)
    assert resid < 0.1
    m.close()


def test_gaussian_sampler_matches_stream():
    """Cheap samples are standard normals of the counter stream, key (seed, 3)."""
    th = OM.init(4).astype(np.float32)
    net = _lib.MLP(th, batch_max=1000)
    n, seed = 3000, 9
    s, s2 = net.gaussian_moments(n, seed)
    z = np.stack([philox.normals(seed, 3, i, [0], 64)[0] for i in range(n)])
    y = net.forward(z.astype(np.float32)).astype(np.float64)
    assert abs(s - y.sum()) < 1e-3 * n and abs(s2 - (y * y).sum()) < 1e-3 * n
    net.close()


def _oracle_at_last_point(q, j):
    """Per-trajectory observables of the oracle path at its last stored point
    (the path runs steps-1 further steps after it, xpath.m:29-86)."""
    snaps = []
    fs = list(q.observe)
    q2 = q.copy()
    q2.observe = [lambda a, p, _f=fs[0]: (snaps.append(a.copy()), _f(a, p))[1]] + fs[1:]
    X.run_path(q2, j, 1)
    a = snaps[-1]
    return np.stack([np.real(f(a, q)).reshape(-1) for f in fs], -1)


@pytest.mark.parametrize("prec,tol,n,lattice", [("fp64", 1e-9, 32, "repulsive"), ("fp32", 2e-4, 32, "repulsive"),
                                               ("fp32", 2e-4, 128, "repulsive"), ("fp32", 2e-4, 128, "attractive")])
def test_tw_trajectory_outputs_match_oracle(prec, tol, n, lattice):
    """Per-trajectory lattice means (|a|^2, |a|^4, low-mode occupation) and the
    noise-mode features against the oracle, trajectory by trajectory; at 32^2 and
    at config 5's 128^2 lattice, with the default repulsive lattice and with the
    attractive lattice of the control-variate benchmark (g = -0.035, hop 270)."""
    kw = dict(g=-0.035, hop=270.0) if lattice == "attractive" else {}
    E = 8 if n == 32 else 2
    p = shrink(tw2d(prec, n=n, modes=32, **kw), E=E, e2=2, points=3, steps=5 if n == 32 else 3)
    p["gpu"] = dict(p["gpu"], observe=["abs2_ave", "abs4_ave", "lowk"])
    prob = Problem(p)
    prob.handle.run_ensemble(1)
    prob.handle.synchronize()
    obs, feat = prob.handle.read_trajectories()
    prob.close()
    q = X.xpreferences(oracle_problem(p))
    modes = lowest_modes(n, n, 32)
    for j in (1, 2):
        ref = _oracle_at_last_point(q, j)                              # [E, 3]
        errs = [_rel(obs[j - 1, 0, :, -1, o], ref[:, o]) for o in range(3)]
        nsp = q.seed + 1 + (j - 1) * q.ensembles[2]
        v = X.xrandom(q, nsp)[0]
        noise = v[0] + 1j * v[1]                                   # [x, y, E]
        F = np.fft.fft2(noise, axes=(0, 1)) * math.sqrt(q.dv / q.nspace)
        fref = np.stack([np.stack([F[kx, ky].real, F[kx, ky].imag], -1) for kx, ky in modes], 1).reshape(-1, 64)
        e3 = np.abs(feat[j - 1, 0] - fref).max()
        print("j=%d |a|^2 %.2e |a|^4 %.2e lowk %.2e features %.2e" % (j, errs[0], errs[1], errs[2], e3))
        # features are stored as float: FP64 lattices agree to float rounding
        assert max(errs[:2]) < tol and e3 < (1e-6 if prec == "fp64" else 2e-3)
        # the low-mode occupation is a small difference of large terms: absolute bound
        assert np.abs(obs[j - 1, 0, :, -1, 2] - ref[:, 2]).max() < (1e-8 if prec == "fp64" else 5e-2)


def _tw_ensembles():
    ptrain = shrink(tw2d("fp32", n=32), E=256, e2=16, points=2, steps=50)
    ptest = dict(ptrain, seed=1000)
    otr, ztr, _ = cv.run_samples(ptrain)
    ote, zte, _ = cv.run_samples(ptest)
    return otr, ztr, ote, zte, cv.feature_translations(ptrain)


def test_control_variate_beats_plain_ensemble():
    """Train on one ensemble of 4096 trajectories, estimate on an independent
    one.  Low-mode occupation: the network alone (no linear part) cuts the
    variance at least 4x, with the same mean within error bars.  g2: never
    worse than the plain ensemble (its fallback), same mean."""
    otr, ztr, ote, zte, tr = _tw_ensembles()
    net = cv.CVModel(seed=0, batch_max=4096)
    info = net.fit(ztr, otr[:, 2], linear=False, translations=tr)
    r0 = cv.estimate(net, zte, ote[:, 2], m_cheap=200000)
    net.close()
    print("lowk, network only", info["beta"], r0)
    assert r0["var_ratio"] > 4.0 and r0["var_ratio_linear"] < 1.5
    assert abs(r0["cv_mean"] - r0["plain_mean"]) < 4 * math.hypot(r0["cv_sigma"], r0["plain_sigma"])

    full = cv.CVModel(seed=0, batch_max=4096)
    info = full.fit(ztr, otr[:, 2], translations=tr)
    r = cv.estimate(full, zte, ote[:, 2], m_cheap=200000)
    full.close()
    print("lowk, linear + network", {k: v for k, v in info.items() if k != "losses"}, r)
    assert 0.0 <= r["beta"] <= 1.0
    assert info["val_ratio"] >= info["val_ratio_linear"] * 0.999
    assert r["var_ratio"] > 4.0 and r["cv_sigma"] < 0.5 * r["plain_sigma"]
    assert abs(r["cv_mean"] - r["plain_mean"]) < 4 * math.hypot(r["cv_sigma"], r["plain_sigma"])

    _, ytr = cv.g2_pseudo(*cv.g2_parts(otr))
    _, yte = cv.g2_pseudo(*cv.g2_parts(ote))
    m = cv.CVModel(seed=0, batch_max=4096)
    info = m.fit(ztr, ytr, translations=tr)
    r = cv.estimate(m, zte, yte, m_cheap=200000)
    m.close()
    print("g2", {k: v for k, v in info.items() if k != "losses"}, r)
    assert r["cv_sigma"] < 1.05 * r["plain_sigma"]
    assert abs(r["cv_mean"] - r["plain_mean"]) < 4 * math.hypot(r["cv_sigma"], r["plain_sigma"])


def test_cross_fitted_estimate():
    """Cross-fitting over two independent folds uses every trajectory: unbiased
    against the pooled plain mean, smaller error bar than the pooled plain one."""
    otr, ztr, ote, zte, tr = _tw_ensembles()
    mu, sig, infos = cv.cross_fit([(ztr, otr[:, 2]), (zte, ote[:, 2])], m_cheap=200000, translations=tr)
    y = np.concatenate([otr[:, 2], ote[:, 2]])
    plain, psig = float(np.mean(y)), float(np.std(y, ddof=1) / math.sqrt(len(y)))
    print("cross-fit %.5f +- %.5f, plain %.5f +- %.5f" % (mu, sig, plain, psig))
    assert abs(mu - plain) < 4 * math.hypot(sig, psig)
    assert sig < 0.5 * psig


@pytest.mark.parametrize("translate", [False, True])
def test_training_graph_matches_eager_epochs(translate):
    """Training epochs replay one captured HIP graph (several epochs in a call)
    or launch eagerly (one epoch per call): same order arrays, translation
    epochs and Adam steps, so the parameters after every call agree bitwise.
    The graph is replayed across two calls with host work and synchronisations
    between the replays."""
    rng = np.random.default_rng(11)
    th = OM.init(3).astype(np.float32)
    x = rng.standard_normal((512, 64)).astype(np.float32)
    t = rng.standard_normal(512).astype(np.float32)
    a, b = _lib.MLP(th, batch_max=1024), _lib.MLP(th, batch_max=1024)
    if translate:
        modes = lowest_modes(16, 16, 32)
        a.set_translations(modes, 16, 16, seed=4)
        b.set_translations(modes, 16, 16, seed=4)
    import ctypes as C
    lib = _lib.load()
    fp = lambda v: v.ctypes.data_as(C.POINTER(C.c_float))
    for call in range(2):
        E = 5
        order = np.concatenate([rng.permutation(512) for _ in range(E)]).astype(np.int32)
        la, lb = np.zeros(E), np.zeros(E)
        _lib.mcheck(lib.xsh_mlp_train(a._h, 512, fp(x), fp(t), E, 64, order.ctypes.data_as(C.POINTER(C.c_int)),
                                      1e-3, la.ctypes.data_as(C.POINTER(C.c_double))))
        for e in range(E):
            oe = np.ascontiguousarray(order[e * 512:(e + 1) * 512])
            l1 = np.zeros(1)
            _lib.mcheck(lib.xsh_mlp_train(b._h, 512, fp(x), fp(t), 1, 64, oe.ctypes.data_as(C.POINTER(C.c_int)),
                                          1e-3, l1.ctypes.data_as(C.POINTER(C.c_double))))
            lb[e] = l1[0]
        pa, pb = a.params(), b.params()
        print("call %d: max |graph - eager| = %.3e, losses %s vs %s" % (call, np.abs(pa - pb).max(), la, lb))
        assert np.array_equal(pa, pb)
        assert np.allclose(la, lb, rtol=1e-6)
    a.close()
    b.close()
