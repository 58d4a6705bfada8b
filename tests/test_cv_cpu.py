"""Host-side estimator logic of the control variate (xspde_matlab_amd/cv.py),
on the CPU: the GPU network is replaced by a stand-in regression with the
same interface, so what is checked is the estimator arithmetic - unbiasedness,
error bars, the ratio pseudo-values, the fallback and cross-fitting - against
synthetic ensembles with known answers."""
import math

import numpy as np
import pytest

from xspde_matlab_amd import _lib, cv


class _StandInNet:
    """Quadratic ridge regression behind the xsh_mlp interface."""

    def __init__(self, theta0, nin=64, nh=256, batch_max=65536):
        self.nin, self.coef = nin, None

    def _feat(self, x):
        x = np.asarray(x, dtype=np.float64)
        return np.c_[np.ones(len(x)), x, x[:, :8] ** 2]

    def set_translations(self, modes, nx, ny, seed=0):
        pass

    def train(self, x, t, epochs, batch, lr, seed=0):
        A = self._feat(x)
        self.coef = np.linalg.solve(A.T @ A + 1e-6 * np.eye(A.shape[1]), A.T @ np.asarray(t, dtype=np.float64))
        return np.zeros(epochs)

    def forward(self, x):
        return (self._feat(x) @ self.coef).astype(np.float32)

    def gaussian_moments(self, n, seed):
        z = np.random.default_rng(seed).standard_normal((n, self.nin))
        y = self._feat(z) @ self.coef
        return float(y.sum()), float((y * y).sum())

    def close(self):
        pass


@pytest.fixture(autouse=True)
def _stand_in(monkeypatch):
    monkeypatch.setattr(_lib, "MLP", _StandInNet)


def _ensemble(n, seed, noise=0.3):
    """y = 2 + z0 + 0.5 (z1^2 - 1) + noise: E[y] = 2 exactly."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((n, 64))
    y = 2.0 + z[:, 0] + 0.5 * (z[:, 1] ** 2 - 1.0) + noise * rng.standard_normal(n)
    return z.astype(np.float32), y


def test_estimate_unbiased_with_smaller_error_bar():
    ztr, ytr = _ensemble(4000, 1)
    zte, yte = _ensemble(4000, 2)
    m = cv.CVModel(seed=0)
    info = m.fit(ztr, ytr, epochs=1)
    r = cv.estimate(m, zte, yte, m_cheap=200000, seed=3)
    assert abs(r["cv_mean"] - 2.0) < 4 * r["cv_sigma"]
    assert abs(r["plain_mean"] - 2.0) < 4 * r["plain_sigma"]
    # var(y) = 1 + 0.5 + 0.09; residual = noise 0.09 (+ fit error)
    assert r["var_ratio"] > 10 and r["cv_sigma"] < 0.4 * r["plain_sigma"]
    assert info["val_ratio"] > 10 and 0.9 < info["beta"] <= 1.0


def test_estimator_z_scores_are_standard():
    """Over repeated independent ensembles the CV error / CV sigma behaves as N(0, 1)."""
    zs = []
    for k in range(40):
        ztr, ytr = _ensemble(800, 100 + k)
        zte, yte = _ensemble(800, 200 + k)
        m = cv.CVModel(seed=k)
        m.fit(ztr, ytr, epochs=1)
        r = cv.estimate(m, zte, yte, m_cheap=50000, seed=300 + k)
        zs.append((r["cv_mean"] - 2.0) / r["cv_sigma"])
    zs = np.array(zs)
    assert abs(zs.mean()) < 0.6 and 0.6 < zs.std() < 1.5


def test_useless_network_falls_back_to_plain():
    rng = np.random.default_rng(5)
    ztr = rng.standard_normal((3000, 64)).astype(np.float32)
    zte = rng.standard_normal((3000, 64)).astype(np.float32)
    ytr, yte = rng.standard_normal(3000), rng.standard_normal(3000)   # independent of z
    m = cv.CVModel(seed=0)
    m.fit(ztr, ytr, epochs=1)
    r = cv.estimate(m, zte, yte, m_cheap=100000)
    assert r["cv_sigma"] < 1.05 * r["plain_sigma"]
    assert abs(r["cv_mean"] - r["plain_mean"]) < 3 * r["plain_sigma"]


def test_g2_pseudo_values():
    rng = np.random.default_rng(7)
    num = 100.0 + rng.standard_normal(5000)
    den = 10.0 + 0.1 * rng.standard_normal(5000) + 0.01 * (num - 100.0)
    g, y = cv.g2_pseudo(num, den)
    assert g == pytest.approx(num.mean() / den.mean() ** 2, rel=1e-14)
    assert y.mean() == pytest.approx(g, rel=1e-12)
    # delta-method error bar against a bootstrap of the ratio
    boot = []
    for b in range(400):
        i = rng.integers(0, 5000, 5000)
        boot.append(num[i].mean() / den[i].mean() ** 2)
    assert np.std(y, ddof=1) / math.sqrt(5000) == pytest.approx(np.std(boot), rel=0.15)


def test_g2_parts_wigner_ordering():
    # symmetric-ordered lattice means <|a|^2> = n, <|a|^4> = n^2 of one trajectory map to
    # <:n^2:> = n^2 - 2n + 1/2 and <n> = n - 1/2 (Np.m:11-16)
    n = 25.0
    obs = np.array([[n, n * n]])
    num, den = cv.g2_parts(obs)
    assert num[0] == pytest.approx(n * n - 2 * n + 0.5)
    assert den[0] == pytest.approx(n - 0.5)


def test_trajectories_to_error_bar():
    assert cv.trajectories_to_error_bar(4.0, 0.0, 10, 0.1) == pytest.approx(400.0)
    assert cv.trajectories_to_error_bar(4.0, 1.0, 50, 0.1) == math.inf    # cheap samples alone too noisy
    assert cv.trajectories_to_error_bar(4.0, 1.0, 10000, 0.1) == pytest.approx(4.0 / (0.01 - 1e-4))


def test_cross_fit():
    z1, y1 = _ensemble(3000, 11)
    z2, y2 = _ensemble(3000, 12)
    mu, sig, infos = cv.cross_fit([(z1, y1), (z2, y2)], m_cheap=200000)
    assert len(infos) == 2
    assert abs(mu - 2.0) < 4 * sig
    plain = np.std(np.r_[y1, y2], ddof=1) / math.sqrt(6000)
    assert sig < 0.4 * plain
