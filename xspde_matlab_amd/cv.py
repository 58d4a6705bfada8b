"""Control-variate estimator for truncated-Wigner ensembles.

The observable of a truncated-Wigner trajectory is a deterministic function of
its initial noise.  A predictor f learns it from the lowest Fourier modes z of
that noise (64 features, i.i.d. standard normal by construction).  For an
independent ensemble of N trajectories and M cheap samples of z,

    mu = mean_N[ Y - f(z) ] + E[f(z)]

is unbiased for E[Y] whatever f is, with variance var(Y - f)/N + var_M/M, where
var_M is the variance of the part of E[f] that is sampled.  The plain ensemble
mean (f = 0) is the fallback.

f has three parts, all fitted on a training ensemble that the estimate does
not use:
  * c0 + c.z, the least-squares linear fit (closed form, FP64 on the host;
    its mean over the feature law is exactly c0, so it needs no sampling);
  * beta * g(z), a 64 -> 256 -> 256 -> 1 GELU network trained on the GPU by the
    xsh_mlp_* kernels (bf16 MFMA, FP32 accumulation, Adam) on the residual of
    the linear fit, with random lattice translations of the feature modes as
    data augmentation (the observables are translation invariant, so each
    trajectory stands for all its translates); its mean is sampled on M cheap
    device-drawn z;
  * beta, the control-variate coefficient cov(r, g)/var(g) of the network on a
    held-out validation split of the training ensemble, clipped to [0, 1], so
    an overfitted network is switched off rather than trusted.

For the ratio g2 = <:n^2:>/<n>^2 the per-trajectory value is the linearised
pseudo-value g2_hat + psi_i (psi the influence function of the ratio), whose
ensemble mean is the plain ratio estimate.
"""
import math
import time

import numpy as np

from . import _lib
from .xsim import Problem


def g2_parts(obs):
    """Per-trajectory numerator and density from [.., nobs] lattice means of
    |a|^2 and |a|^4 in Wigner (symmetric) ordering
    (Applications/Phase-space/GBSObserve/Np.m:11-16): <:n^2:> = <|a|^4> - 2<|a|^2> + 1/2,
    <n> = <|a|^2> - 1/2; g2 = <:n^2:>/<n>^2."""
    a2, a4 = obs[..., 0], obs[..., 1]
    return a4 - 2.0 * a2 + 0.5, a2 - 0.5


def g2_pseudo(num, den):
    """Plain g2 and per-trajectory pseudo-values whose mean is that g2 and whose
    spread is the first-order (delta-method) fluctuation of the ratio."""
    mn, md = float(np.mean(num)), float(np.mean(den))
    g = mn / (md * md)
    psi = (np.asarray(num) - mn) / (md * md) - 2.0 * g * (np.asarray(den) - md) / md
    return g, g + psi


def run_samples(p, npar=1):
    """Simulate an ensemble; returns (observables [N, nobserve] at the final
    output point, features [N, 2 nmodes], kernel seconds), one row per trajectory."""
    prob = Problem(p)
    try:
        prob.handle.run_ensemble(npar)
        ms = prob.handle.synchronize()
        obs, feat = prob.handle.read_trajectories()
    finally:
        prob.close()
    o = obs[:, 0, :, -1, :]                                     # [e2, E, nobserve]
    return o.reshape(-1, o.shape[-1]), feat[:, 0].reshape(-1, feat.shape[-1]), ms / 1e3


def feature_translations(p):
    """(modes, nx, ny) of a problem's feature modes, for CVModel.fit."""
    from .xsim import lowest_modes
    nx, ny = p["points"][1], p["points"][2] if len(p["points"]) > 2 else 1
    m = p["gpu"]["modes"]
    return (lowest_modes(nx, ny, m) if isinstance(m, int) else [tuple(t) for t in m]), nx, ny


class CVModel:
    def __init__(self, nin=64, nh=256, seed=0, batch_max=65536):
        theta = _init_params(nin, nh, seed)
        self.mlp = _lib.MLP(theta, nin, nh, batch_max)
        self.nin = nin
        self.c0, self.c = 0.0, np.zeros(nin)
        self.mean, self.std, self.beta = 0.0, 1.0, 0.0

    def fit(self, z, y, epochs=300, batch=512, lr=2e-3, seed=0, val_frac=0.25, linear=True, translations=None):
        """Fit the predictor on (z, y); returns a dict with the per-epoch loss,
        beta and validation variance ratios (plain/linear, plain/full).

        Minibatches of 512 at lr 2e-3: a quarter of the optimizer steps of
        batch 128 at lr 1e-3 (each step is a chain of small kernels whose
        latency, not their flops, sets the time), the same g2 residual variance
        on the 128x128 truncated-Wigner workload (ratio 17.2 vs 17.4; low-mode
        occupation 46.7 vs 54.0), fit 0.29 s instead of 0.69 s.

        ``translations = (modes, nx, ny)`` declares the features to be the
        Fourier modes ``modes`` of an nx x ny periodic lattice and y to be
        translation invariant: the network then trains on randomly translated
        copies of the samples (a fresh lattice shift per sample and epoch)."""
        z = np.asarray(z, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        n = len(y)
        nv = int(round(val_frac * n)) if val_frac > 0 else 0
        perm = np.random.default_rng(seed + 17).permutation(n)
        tr, va = perm[nv:], perm[:nv]
        if linear:
            # a translation-invariant observable can only depend linearly on
            # the k = 0 mode; fitting the other coefficients would fit noise
            cols = np.arange(self.nin)
            if translations is not None:
                k0 = [m for m, (kx, ky) in enumerate(translations[0]) if kx == 0 and ky == 0]
                cols = np.array([c for m in k0 for c in (2 * m, 2 * m + 1)], dtype=int)
            A = np.c_[np.ones(len(tr)), z[tr][:, cols]]
            coef, *_ = np.linalg.lstsq(A, y[tr], rcond=None)
            self.c0, self.c = float(coef[0]), np.zeros(self.nin)
            self.c[cols] = coef[1:]
        else:
            self.c0, self.c = float(np.mean(y[tr])), np.zeros(self.nin)
        r = y - self.c0 - z @ self.c
        self.mean, self.std = float(np.mean(r[tr])), float(np.std(r[tr])) or 1.0
        if translations is not None:
            modes, nx, ny = translations
            self.mlp.set_translations(modes, nx, ny, seed)
        else:
            self.mlp.set_translations(None, 1, 1)
        losses = self.mlp.train(z[tr].astype(np.float32), ((r[tr] - self.mean) / self.std).astype(np.float32),
                                epochs, min(batch, len(tr)), lr, seed)
        out = dict(losses=losses, beta=1.0)
        if nv > 1:
            g = self._net(z[va])
            vg = np.var(g)
            beta = float(np.cov(r[va], g)[0, 1] / vg) if vg > 0 else 0.0
            self.beta = min(max(beta, 0.0), 1.0)
            vy = np.var(y[va])
            out.update(beta=self.beta, val_ratio_linear=vy / max(np.var(r[va]), 1e-300),
                       val_ratio=vy / max(np.var(r[va] - self.beta * g), 1e-300))
        else:
            self.beta = 1.0
        return out

    def _net(self, z):
        return self.mlp.forward(np.asarray(z, dtype=np.float32)).astype(np.float64) * self.std + self.mean

    def linear(self, z):
        return self.c0 + np.asarray(z, dtype=np.float64) @ self.c

    def predict(self, z):
        f = self.linear(z)
        if self.beta:
            f = f + self.beta * self._net(z)
        return f

    def cheap_moments(self, m, seed):
        """E[f] estimated over m device-drawn feature vectors, and the variance
        of the sampled part (the linear part's mean c0 is exact)."""
        if not self.beta:
            return self.c0, 0.0
        s, s2 = self.mlp.gaussian_moments(m, seed)
        mu = s / m
        var = max(s2 / m - mu * mu, 0.0)
        b = self.beta
        return self.c0 + b * (mu * self.std + self.mean), (b * self.std) ** 2 * var

    def close(self):
        self.mlp.close()


def _init_params(nin, nh, seed):
    """He initialisation, flat layout W1 b1 W2 b2 w3 b3."""
    rng = np.random.default_rng(seed)
    return np.concatenate([
        rng.standard_normal(nin * nh) * math.sqrt(2.0 / nin), np.zeros(nh),
        rng.standard_normal(nh * nh) * math.sqrt(2.0 / nh), np.zeros(nh),
        rng.standard_normal(nh) * math.sqrt(1.0 / nh), np.zeros(1)]).astype(np.float32)


def estimate(model, z, y, m_cheap=10 ** 6, seed=7):
    """Control-variate, linear-only and plain estimates of E[y] with standard errors
    and the residual variance ratios var(y)/var(y - f)."""
    y = np.asarray(y, dtype=np.float64)
    n = len(y)
    f = model.predict(z)
    d = y - f
    mu_f, var_f = model.cheap_moments(m_cheap, seed)
    dl = y - model.linear(z)
    vy = np.var(y, ddof=1)
    return dict(cv_mean=float(np.mean(d) + mu_f),
                cv_sigma=math.sqrt(np.var(d, ddof=1) / n + var_f / m_cheap),
                linear_mean=float(np.mean(dl) + model.c0),
                linear_sigma=math.sqrt(np.var(dl, ddof=1) / n),
                plain_mean=float(np.mean(y)), plain_sigma=math.sqrt(vy / n),
                var_ratio=float(vy / max(np.var(d, ddof=1), 1e-300)),
                var_ratio_linear=float(vy / max(np.var(dl, ddof=1), 1e-300)),
                var_cheap=var_f, beta=model.beta, n=n, m_cheap=m_cheap)


def cross_fit(folds, m_cheap=10 ** 6, seed=7, batch_max=8192, **fit_kw):
    """Cross-fitted control-variate estimate over independent folds [(z, y), ...]:
    the model for fold k is trained on the other folds, so every trajectory
    enters the estimate and each prediction is independent of its own
    trajectory.  Returns (mean, sigma, per-fold fit info)."""
    n = sum(len(y) for _, y in folds)
    total, var, infos = 0.0, 0.0, []
    for k, (zk, yk) in enumerate(folds):
        zt = np.concatenate([z for j, (z, _) in enumerate(folds) if j != k])
        yt = np.concatenate([y for j, (_, y) in enumerate(folds) if j != k])
        m = CVModel(seed=k, batch_max=batch_max)
        info = m.fit(zt, yt, **fit_kw)
        d = np.asarray(yk, dtype=np.float64) - m.predict(zk)
        mu_f, var_f = m.cheap_moments(m_cheap, seed + k)
        m.close()
        w = len(yk) / n
        total += w * (float(np.mean(d)) + mu_f)
        var += w * w * (np.var(d, ddof=1) / len(yk) + var_f / m_cheap)
        infos.append(info)
    return total, math.sqrt(var), infos


def trajectories_to_error_bar(var_resid, var_f, m_cheap, sigma_target):
    """Test trajectories the control-variate estimate needs to reach sigma_target."""
    budget = sigma_target ** 2 - var_f / m_cheap
    return var_resid / budget if budget > 0 else float("inf")


def timed(fn, *a, **k):
    t0 = time.perf_counter()
    r = fn(*a, **k)
    return r, time.perf_counter() - t0
