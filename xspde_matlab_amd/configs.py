"""The benchmark workloads as xSPDE project dictionaries (built-in GPU forms).

Each is an xSPDE project dictionary (the reference's field names) with the
built-in GPU forms in ``gpu``.  Defaults are the full workload sizes;
``shrink`` makes smaller copies for tests.
"""
import math


def kubo(precision="fp64", E=4096, e2=16, points=201):
    # Kubo oscillator: da = i a xi dt, a(0)=1, MP, dt=0.01 to t=2 (Kubo.m:12-19)
    return dict(name="Kubo", dimensions=1, points=[points], ranges=[2.0], steps=[1], checks=1,
                noises=1, method="MP", ensembles=[E // e2, e2, 1],
                gpu=dict(deriv="kubo", c=1.0, initial="const", c0=1.0, observe=["a", "abs2"],
                         precision=precision))


def nls1d(precision="fp32", E=4096, e2=16, points=101, nx=256, steps=10, gamma=0.1, eps=0.1):
    # 1D stochastic NLS / GP: da = [i/2 d2a - i|a|^2 a - gamma a] dt + sqrt(gamma) dW
    return dict(name="NLS1D", dimensions=2, points=[points, nx], ranges=[5.0, 20.0], steps=[steps, 1],
                checks=1, noises=2, inrandoms=2, method="RK4", ensembles=[E // e2, e2, 1],
                gpu=dict(deriv="gp", g=-1j, u=-gamma, s=math.sqrt(gamma / 2.0),
                         linear=dict(l2=0.5j), initial="sech", amp=1.0, width=1.0, eps=eps,
                         observe=["abs2", "a"], precision=precision))


def cgl2d(precision="fp32", E=1024, e2=16, points=101, n=128, steps=20, length=20.0):
    # 2D damped CGL: da = [(1+i) lap a + a - (1+0.5i)|a|^2 a] dt + 0.05 dW (complex)
    return dict(name="CGL2D", dimensions=3, points=[points, n, n], ranges=[4.0, length, length],
                steps=[steps, 1, 1], checks=1, noises=2, urandoms=1, method="RK4",
                ensembles=[E // e2, e2, 1],
                gpu=dict(deriv="gp", g=-(1 + 0.5j), u=0.0, s=0.05 / math.sqrt(2.0),
                         linear=dict(l0=1.0, l2=1 + 1j), initial="phase", amp=0.1,
                         observe=["abs2_ave", "corr"], precision=precision))


def shrink(p, E=None, e2=None, points=None, steps=None):
    p = dict(p)
    p["gpu"] = dict(p["gpu"])
    ens = list(p["ensembles"])
    if E is not None:
        ens[0] = E
    if e2 is not None:
        ens[1] = e2
    p["ensembles"] = ens
    # keep the coarse time step ranges(1)/((points(1)-1)*steps(1)) unchanged
    dtc = p["ranges"][0] / ((p["points"][0] - 1) * p["steps"][0])
    if points is not None:
        p["points"] = [points] + list(p["points"][1:])
    if steps is not None:
        p["steps"] = [steps] + list(p["steps"][1:])
    p["ranges"] = [dtc * p["steps"][0] * (p["points"][0] - 1)] + list(p["ranges"][1:])
    return p


def tw2d(precision="fp32", E=1024, e2=64, n=128, steps=100, points=11, g=0.01, n0=100.0, modes=32, seed=0,
         hop=0.5):
    """Truncated-Wigner 2D Bose lattice GP: a(0) = sqrt(n0) + Wigner vacuum noise
    (variance 1/2 per mode on a unit-spacing lattice), da = [i hop lap a
    - i g (|a|^2 - 1) a] dt, RK4 interaction picture, dt = 0.001 to t = 1.
    Observables: lattice means of |a|^2 and |a|^4 (g2 numerator and
    denominator) and the occupation of the nonzero feature modes, kept per
    trajectory with the lowest Fourier modes of the initial noise (features of
    the control variate)."""
    return dict(name="TW2D", dimensions=3, points=[points, n, n], ranges=[(points - 1) * steps * 0.001, n - 1, n - 1],
                steps=[steps, 1, 1], checks=0, noises=0, inrandoms=2, method="RK4", ensembles=[E, e2, 1], seed=seed,
                gpu=dict(deriv="gp", g=-1j * g, u=1j * g, s=0.0, linear=dict(l2=1j * hop), initial="coherent",
                         alpha=n0 ** 0.5, observe=["abs2_ave", "abs4_ave", "lowk"], precision=precision,
                         traj_out=True, modes=modes))


BATCH_1D = (32, 64, 128, 256, 512)
BATCH_2D = (32, 64, 128)
BATCH_E = (256, 512, 1024, 2048, 4096)


def batch_instances(count=64, seed=0, steps=500, precision="fp32", points=11, sizes_1d=BATCH_1D,
                    sizes_2d=BATCH_2D):
    """Uneven batch: `count` independent GP instances, lattice drawn by seed from
    1D {32..512} and 2D {32^2, 64^2, 128^2}, damping gamma ~ U[0.01, 0.5],
    nonlinearity g ~ U[0.1, 2], ensemble size from {256..4096}, `steps` RK4IP
    steps (points-1 stored intervals), sech/plane-wave initial state plus noise:
        da = [i/2 lap a - i g |a|^2 a - gamma a] dt + sqrt(gamma) dW."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = [(n,) for n in sizes_1d] + [(n, n) for n in sizes_2d]
    per = steps // (points - 1)
    out = []
    for i in range(count):
        shape = sizes[rng.integers(len(sizes))]
        gamma = float(rng.uniform(0.01, 0.5))
        g = float(rng.uniform(0.1, 2.0))
        E = int(BATCH_E[rng.integers(len(BATCH_E))])
        e2 = 4 if E >= 1024 else 1
        nd = len(shape) + 1
        length = 20.0
        p = dict(name="B%02d" % i, dimensions=nd, points=[points] + list(shape),
                 ranges=[0.005 * steps] + [length] * len(shape), steps=[per] + [1] * len(shape), checks=0,
                 noises=2, inrandoms=2, method="RK4", ensembles=[E // e2, e2, 1], seed=i,
                 gpu=dict(deriv="gp", g=-1j * g, u=-gamma, s=math.sqrt(gamma / 2.0),
                          linear=dict(l2=0.5j), initial="sech", amp=1.0, width=1.0, eps=0.1,
                          observe=["abs2_ave", "a_ave"], precision=precision))
        out.append(p)
    return out
