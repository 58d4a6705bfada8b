"""GPU parity of the HIP path against the FP64 oracle, through the C ABI.

Same seeded counter stream on both sides, so trajectories are compared
directly.  Tolerances: FP64 GPU vs FP64 oracle 1e-9 relative (transcendental
rounding of the noise transform only); FP32 GPU vs FP64 oracle as stated per
test (single-precision rounding, grows with the number of steps).
"""
import numpy as np
import pytest

from oracle import philox
from oracle import xspde as X
from xspde_matlab_amd import Problem
from xspde_matlab_amd.xsim import preferences

from _problems import cgl2d, kubo, nls1d, oracle_problem, shrink

pytestmark = pytest.mark.gpu


def _oracle_path(p, j, nc, npar=1):
    q = X.xpreferences(oracle_problem(p))
    a, av = X.run_path(q, j, nc, npar)
    return q, a, av


def _cmp(gpu, ref, rtol, what):
    gpu, ref = np.asarray(gpu), np.asarray(ref)
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(gpu - ref).max() / scale
    print("%s: max |gpu-oracle|/max|oracle| = %.3e (tol %.1e)" % (what, err, rtol), flush=True)
    assert err <= rtol, what


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
def test_normals_match_oracle(prec, tol):
    prob = Problem(shrink(kubo(prec), E=8, e2=1, points=3))
    for tag, draw in [(philox.TAG_NOISE, 0), (philox.TAG_NOISE, 17), (philox.TAG_GAUSS_INIT, 0)]:
        g = prob.handle.probe_normals(5, tag, draw, 3, 7, 37)
        r = philox.normals(5, tag, draw, np.arange(3, 10), 37)
        _cmp(g, r, tol, "normals %s tag %d draw %d" % (prec, tag, draw))
    prob.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 3e-6)])
@pytest.mark.parametrize("shape", [(32,), (64,), (128,), (256,), (512,), (1024,), (32, 32), (64, 64), (128, 128)])
def test_propagator_matches_oracle(prec, tol, shape):
    nd = len(shape) + 1
    p = dict(dimensions=nd, points=[3] + list(shape), ranges=[1.0] + [10.0] * len(shape), ensembles=[2, 1, 1],
             noises=0, gpu=dict(deriv="zero", linear=dict(l0=0.3 - 0.2j, l1=0.1 + 0.05j, l2=0.4 + 0.9j),
                                initial="const", observe=["a"], precision=prec))
    prob = Problem(p)
    rng = np.random.default_rng(7)
    planes = rng.standard_normal((3,) + shape[::-1]) + 1j * rng.standard_normal((3,) + shape[::-1])
    got = prob.handle.probe_propagate(1, planes.reshape(3, -1)).reshape(planes.shape)
    q = X.xpreferences(oracle_problem(p))
    q.dtr = q.dt / 2
    prop = X.xpropfactor(q)
    a = np.moveaxis(planes, 0, -1)            # [y, x, plane] -> need [1, x, (y), plane]
    a = np.transpose(a, tuple(range(len(shape) - 1, -1, -1)) + (len(shape),))[None]
    ref = X.xprop(a, q, prop)[0]
    ref = np.moveaxis(np.transpose(ref, tuple(range(len(shape) - 1, -1, -1)) + (len(shape),)), -1, 0)
    _cmp(got, ref, tol, "propagate %s %s" % (prec, shape))
    prob.close()


def _path_case(p, tol, nc_list=(1, 2), j=2):
    prob = Problem(p)
    for nc in nc_list:
        av_g, a_g = prob.xpath(nc, npar=1, j=j, want_final=True)
        q, a_r, av_r = _oracle_path(p, j, nc)
        # final fields: oracle [1, x, (y), E] -> [E, site] with site = ix + nx*iy
        ar = np.moveaxis(a_r[0], -1, 0)
        ar = ar.reshape(ar.shape[0], -1, order="F") if ar.ndim > 2 else ar.reshape(ar.shape[0], -1)
        _cmp(a_g.reshape(ar.shape), ar, tol, "%s final field nc=%d" % (p["name"], nc))
        for o, (g, r) in enumerate(zip(av_g, av_r)):
            if r.shape[2] > 1 and len(q.space) == 2:
                r = r.reshape(r.shape[0], r.shape[1], q.space[0], q.space[1]).reshape(
                    r.shape[0], r.shape[1], -1, order="F")
            _cmp(g, r, tol, "%s av[%d] nc=%d" % (p["name"], o, nc))
    prob.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 2e-5)])
@pytest.mark.parametrize("method", ["MP", "RK4"])
def test_kubo_path(prec, tol, method):
    p = shrink(kubo(prec), E=300, e2=2, points=11)
    p["method"] = method
    _path_case(p, tol)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 5e-5)])
@pytest.mark.parametrize("nx", [64, 256])
@pytest.mark.parametrize("method", ["RK4", "MP"])
def test_nls1d_path(prec, tol, nx, method):
    p = shrink(nls1d(prec, nx=nx), E=20, e2=2, points=4, steps=3)
    p["method"] = method
    _path_case(p, tol)


@pytest.mark.parametrize("prec,tol,n", [("fp64", 1e-10, 32), ("fp64", 1e-10, 64), ("fp64", 1e-10, 128),
                                        ("fp32", 5e-5, 32), ("fp32", 5e-5, 64), ("fp32", 5e-5, 128)])
@pytest.mark.parametrize("method", ["RK4", "MP"])
def test_cgl2d_path(prec, tol, n, method):
    p = shrink(cgl2d(prec, n=n), E=3, e2=2, points=3, steps=2)
    p["method"] = method
    _path_case(p, tol)


def test_kubo_mean_within_error_bars():
    """<Re a>(t) = exp(-t/2) (Kubo.m:19) within 4 sampling error bars, 4096 trajectories."""
    from xspde_matlab_amd import xsim
    p = kubo("fp64")
    err, data, q, _ = xsim(p)
    t = np.linspace(0, 2, q["points"][0])
    mean, step, sig = data[0][:, 0], data[0][:, 1], data[0][:, 2]
    dev = np.abs(mean - np.exp(-t / 2))
    print("kubo: max dev %.3e, max sigma %.3e, max step %.3e" % (dev.max(), sig.max(), step.max()))
    assert np.all(dev <= 4 * sig + 2 * step + 1e-12)
    assert np.all(np.abs(data[1][:, 0] - 1) <= 4 * data[1][:, 2] + 2 * data[1][:, 1] + 1e-6)


@pytest.mark.parametrize("order,e2,e3", [(0, 3, 2), (2, 4, 1), (1, 2, 3)])
def test_ensemble_matches_oracle_fp64(order, e2, e3):
    """xsim on the GPU (one launch of all (j, nc) paths per parallel ensemble, the
    host's D/D^2 bookkeeping and statistics) against the oracle's xsim: every
    column of every output (mean, step error, sampling error) and the error
    vector (total, step, sampling), for Richardson orders 0..2 and e2 e3 > 1."""
    from xspde_matlab_amd import xsim
    p = shrink(nls1d("fp64", nx=64), E=8, e2=e2, points=4, steps=2)
    p["ensembles"][2] = e3
    p["order"] = order
    err_g, data_g, q, _ = xsim(p)
    err_r, outd, qo = X.xsim(oracle_problem(p))
    sx = max(3, 1 + qo.nchk)
    for o in range(len(outd)):
        ref = outd[o].reshape(-1, outd[o].shape[-1])
        _cmp(data_g[o][:, 0], ref[:, 0], 1e-10, "ensemble mean o=%d" % o)
        _cmp(data_g[o][:, 1], ref[:, 1], 1e-8, "step error o=%d" % o)
        _cmp(data_g[o][:, 2], ref[:, sx - 1], 1e-8, "sampling error o=%d" % o)
    print("error vector gpu", err_g[:sx + 1], "oracle", err_r[:sx + 1])
    assert np.allclose(err_g[:sx + 1], err_r[:sx + 1], rtol=1e-8, atol=0)


BOUNDS = {"NN": [-1, -1], "DD": [1, 1], "DN": [1, -1], "ND": [-1, 1], "PP": [0, 0]}


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 3e-6)])
@pytest.mark.parametrize("shape,bnd", [((64,), ("NN",)), ((128,), ("DD",)), ((128,), ("DN",)), ((64,), ("ND",)),
                                       ((256,), ("NN",)), ((64, 64), ("NN", "NN")), ((64, 64), ("PP", "NN")),
                                       ((64, 64), ("DD", "ND")), ((128, 128), ("NN", "NN"))])
def test_boundary_propagator_matches_oracle(prec, tol, shape, bnd):
    """Dense DCT/DST (and DFT on mixed lattices) propagators against xprop/xtransform
    (xtransform.m:33-70) for every boundary type of spdpreferences.m:56-67."""
    nd = len(shape) + 1
    b = [BOUNDS[x] for x in bnd]
    p = dict(dimensions=nd, points=[3] + list(shape), ranges=[1.0] + [10.0] * len(shape), ensembles=[2, 1, 1],
             noises=0, boundaries=b,
             gpu=dict(deriv="zero", linear=dict(l0=0.3 - 0.2j, l2=0.4 + 0.9j), initial="const",
                      observe=["a"], precision=prec))
    prob = Problem(p)
    rng = np.random.default_rng(11)
    planes = rng.standard_normal((3,) + shape[::-1]) + 1j * rng.standard_normal((3,) + shape[::-1])
    got = prob.handle.probe_propagate(1, planes.reshape(3, -1)).reshape(planes.shape)
    q = X.xpreferences(oracle_problem(p))
    q.dtr = q.dt / 2
    prop = X.xpropfactor(q)
    a = np.moveaxis(planes, 0, -1)
    a = np.transpose(a, tuple(range(len(shape) - 1, -1, -1)) + (len(shape),))[None]
    ref = X.xprop(a, q, prop)[0]
    ref = np.moveaxis(np.transpose(ref, tuple(range(len(shape) - 1, -1, -1)) + (len(shape),)), -1, 0)
    _cmp(got, ref, tol, "propagate %s %s %s" % (prec, shape, bnd))
    prob.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 5e-5)])
@pytest.mark.parametrize("method", ["RK4", "MP"])
@pytest.mark.parametrize("bnd", ["NN", "DD", "DN"])
def test_nls1d_boundary_path(prec, tol, method, bnd):
    """Whole paths with non-periodic ends, including xsetbound's Dirichlet values."""
    p = shrink(nls1d(prec, nx=128), E=64, e2=2, points=3, steps=2)
    p["method"] = method
    p["boundaries"] = [BOUNDS[bnd]]
    _path_case(p, tol)


@pytest.mark.parametrize("prec,tol,n", [("fp64", 1e-10, 64), ("fp32", 5e-5, 64), ("fp32", 5e-5, 128)])
def test_cgl2d_neumann_path(prec, tol, n):
    """Config 3-NN: 2D CGL with Neumann ends (DCT-I both dimensions)."""
    p = shrink(cgl2d(prec, n=n), E=2, e2=2, points=3, steps=2)
    p["boundaries"] = [[-1, -1], [-1, -1]]
    # corr is the periodic FFT correlation of the field whatever the boundaries
    p["gpu"]["observe"] = ["abs2_ave", "abs2", "corr"]
    _path_case(p, tol)


@pytest.mark.parametrize("method", ["RK4", "MP"])
def test_stream_kernel_persistent_grid(method):
    """FP64 128x128 runs on the streaming kernel, whose persistent grid holds
    at most 1024 workgroups: with 1100 trajectories the last 76 are integrated
    by workgroups on their second work item.  Trajectories depend only on their
    own counter stream, so the first three must equal a 3-trajectory run
    bit for bit, and the ensemble mean of |a|^2 at the last point must equal
    the mean over the returned final fields of all 1100 trajectories."""
    base = shrink(cgl2d("fp64", n=128), E=3, e2=1, points=2, steps=1)
    base["method"] = method
    base["gpu"] = dict(base["gpu"], observe=["abs2_ave"])
    prob = Problem(base)
    assert prob.info.threads == 1024 and prob.info.traj_per_block == 1
    av3, a3 = prob.xpath(1, npar=1, j=1, want_final=True)
    prob.close()
    big = dict(base, ensembles=[1100, 1, 1])
    prob = Problem(big)
    av, a = prob.xpath(1, npar=1, j=1, want_final=True)
    prob.close()
    a = np.asarray(a).reshape(1100, -1)
    assert np.array_equal(a[:3], np.asarray(a3).reshape(3, -1))
    m = np.mean(np.abs(a) ** 2)
    print("ensemble |a|^2 %.15g, from final fields %.15g" % (av[0][0, -1, 0].real, m))
    assert abs(av[0][0, -1, 0].real - m) < 1e-12 * m


@pytest.mark.parametrize("case", ["nls1d-fp64-256", "nls1d-fp32-64", "cgl2d-fp64-64", "cgl2d-fp32-128",
                                  "cgl2d-fp64-128", "cgl2d-nn-fp32-128"])
def test_kspace_observable(case):
    """|a_k|^2 per k site (the observable |a|^2 with p.transforms on every space
    dimension, xnft.m): accumulated inside the transform's middle operation, for
    partial trajectory groups (20 trajectories, 16 per workgroup in 1D), the
    streaming FP64 kernel and a Neumann lattice (the FFT of the field whatever
    the boundaries)."""
    kind, *rest = case.split("-")
    if kind == "nls1d":
        prec, n = rest[0], int(rest[1])
        p = shrink(nls1d(prec, nx=n), E=20, e2=2, points=3, steps=2)
    else:
        nn = rest[0] == "nn"
        prec, n = rest[-2], int(rest[-1])
        p = shrink(cgl2d(prec, n=n), E=3, e2=2, points=3, steps=2)
        if nn:
            p["boundaries"] = [[-1, -1], [-1, -1]]
    p["gpu"]["observe"] = ["abs2k", "abs2_ave"]
    _path_case(p, 1e-10 if prec == "fp64" else 5e-5)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 5e-5)])
@pytest.mark.parametrize("bnd", ["NN", "DD", "DN", "ND"])
def test_nls1d_boundary_values(prec, tol, bnd):
    """Constant non-zero boundary values (p.boundval): the boundary patch around
    every propagation, the Neumann correction and the Dirichlet values, against
    the oracle (whose patches reproduce the inhomogeneous heat-equation closed
    forms)."""
    code = {"NN": [-1, -1], "DD": [1, 1], "DN": [1, -1], "ND": [-1, 1]}[bnd]
    p = shrink(nls1d(prec, nx=128), E=20, e2=2, points=3, steps=2)
    p["boundaries"] = [code]
    p["boundval"] = [[0.3 + 0.1j, -0.2 + 0.05j]]
    p["gpu"]["observe"] = ["abs2", "a"]
    _path_case(p, tol)


@pytest.mark.parametrize("prec,tol,n", [("fp64", 1e-10, 64), ("fp32", 5e-5, 128)])
def test_cgl2d_boundary_values(prec, tol, n):
    """2D: Neumann in x with derivative values, Dirichlet in y with field values."""
    p = shrink(cgl2d(prec, n=n), E=2, e2=2, points=3, steps=2)
    p["boundaries"] = [[-1, -1], [1, 1]]
    p["boundval"] = [[0.02, -0.01 + 0.005j], [0.05 - 0.02j, 0.03]]
    # the lattice mean of a is a near-cancelling sum (random phases): FP64 only
    p["gpu"]["observe"] = ["abs2_ave", "abs2"] + (["a_ave"] if prec == "fp64" else [])
    _path_case(p, tol)


def _ends_1d(a, c, d, p):
    """Time-dependent boundary values of the x ends (p.boundfun)."""
    return (0.3 * np.exp(2j * p.t) + 0.1 * p.t, (-0.2 + 0.05j) * np.cos(3 * p.t))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 5e-5)])
@pytest.mark.parametrize("method", ["RK4", "MP"])
@pytest.mark.parametrize("bnd", ["NN", "DD", "DN", "ND"])
def test_nls1d_time_dependent_boundaries(prec, tol, method, bnd):
    """p.boundfun values changing in time: tabulated per half fine step on the
    host, read by the patch, the Neumann correction and the Dirichlet ends at the
    method's p.t of each stage (start, midpoint, end; coarse steps at every
    other level), against the oracle (pinned on Batchtest #42)."""
    p = shrink(nls1d(prec, nx=128), E=20, e2=2, points=3, steps=2)
    p["method"] = method
    p["boundaries"] = [BOUNDS[bnd]]
    p["boundfun"] = _ends_1d
    p["gpu"]["observe"] = ["abs2", "a"]
    _path_case(p, tol)


def _ends_2d(a, c, d, p):
    """x ends vary along y, y ends along x, both in time."""
    if d == 2:
        return (0.02 * np.cos(p.y) * np.exp(1j * p.t), 0.01 * np.sin(p.y + p.t))
    return (0.05 * np.exp(-p.t) * np.cos(0.5 * p.x), 0.03 + 0.01j * p.t)


@pytest.mark.parametrize("prec,tol,n", [("fp64", 1e-10, 64), ("fp32", 5e-5, 128), ("fp64", 1e-10, (48, 40)),
                                        ("fp32", 5e-5, (40, 48))])
def test_cgl2d_time_dependent_boundaries(prec, tol, n):
    """2D: values along the boundary lines and in time (Neumann x, Dirichlet y);
    also on padded lattices (48 x 40, 40 x 48 on the 64 x 64 kernel), where the
    caller's table rows are max(nx, ny) wide and the kernel's 64."""
    p = shrink(cgl2d(prec, n=n if isinstance(n, int) else 32), E=2, e2=2, points=3, steps=2)
    if not isinstance(n, int):
        p["points"] = [p["points"][0], n[0], n[1]]
    p["boundaries"] = [[-1, -1], [1, 1]]
    p["boundfun"] = _ends_2d
    p["gpu"]["observe"] = ["abs2_ave", "abs2"] + (["a_ave"] if prec == "fp64" else [])
    _path_case(p, tol)


def _peregrine(x, t):
    return np.exp(1j * t) * (4 * (1 + 2j * t) / (1 + 4 * (t ** 2 + x ** 2)) - 1)


def _peregrine_dx(x, t):
    return -8 * x * np.exp(1j * t) * (4 * (1 + 2j * t) / (1 + 4 * (t ** 2 + x ** 2)) ** 2)


_PEREGRINE_ENDS = {
    "DD": lambda x0, t: (_peregrine(x0, t), _peregrine(x0, t)),
    "NN": lambda x0, t: (_peregrine_dx(x0, t), -_peregrine_dx(x0, t)),
    "DN": lambda x0, t: (_peregrine(x0, t), -_peregrine_dx(x0, t)),
    "ND": lambda x0, t: (_peregrine_dx(x0, t), _peregrine(x0, t)),
}


@pytest.mark.parametrize("prec,points,tol,etol", [("fp64", 51, 1e-8, 1e-4), ("fp32", 11, 2e-4, None)])
def test_peregrine_time_dependent_boundaries(prec, points, tol, etol):
    """Examples/SPDE_2D/Peregrine.m on a 128-site lattice (the GPU lattices are
    even; the oracle reproduces the reference's 161-site Batchtest #42 value):
    NLS with the Peregrine solution's values at both ends for each boundary type,
    MP with checks and order-2 extrapolation.  Every (nc) path against the oracle,
    and the xsim error vector (step error, difference from the exact solution)
    of the GPU averages against the oracle's.  The Peregrine wave sits on a
    modulationally unstable plane-wave background, so rounding differences grow
    about as e^t (a 1e-15 change of the initial field reaches 7e-12 at t = 5 in
    the oracle): FP64 runs the reference's whole t = -5..5 (3057 steps) at 1e-8,
    FP32 the first two time units (t = -5..-3, 657 steps) at 2e-4.  The error
    vector is compared in FP64 (to 1e-4 of each entry); FP32 rounding is of the
    size of the short run's errors themselves."""
    ends = {"DD": [1, 1], "NN": [-1, -1], "DN": [1, -1], "ND": [-1, 1]}
    datas_g, datas_r = [], []
    for name, b in ends.items():
        f = _PEREGRINE_ENDS[name]
        p = dict(name="Peregrine-" + name, dimensions=2, noises=1, ranges=[0.2 * (points - 1), 10.0],
                 origins=[-5.0, -5.0],
                 points=[points, 128], method="MP", order=2, steps=20, checks=1, ensembles=[1, 1, 1],
                 boundaries=[b], boundfun=lambda a, c, d, p, _f=f: _f(p.origins[1], p.t),
                 gpu=dict(deriv="gp", g=1j, s=0.0, linear=dict(l2=0.5j), initial="const", c0=0.0,
                          observe=["a"], precision=prec))
        q = X.xpreferences(oracle_problem(p))
        q.output = [lambda o, p: np.abs(o[0]) ** 2]
        q.compare = [lambda p: np.abs(_peregrine(p.x, p.t)) ** 2]
        x = q.xc[1]
        a0 = _peregrine(x, -5.0)
        prob = Problem(p)
        gpu_av = {}
        for nc in (1, 2):
            av_g, a_g = prob.xpath(nc, npar=1, j=1, a0=a0[None, :], want_final=True)
            a_r, av_r = X.run_path(q, 1, nc, a0=a0.reshape(1, -1, 1))
            _cmp(a_g.reshape(-1), a_r.reshape(-1), tol, "Peregrine %s final field nc=%d" % (name, nc))
            _cmp(av_g[0], av_r[0], tol, "Peregrine %s <a> nc=%d" % (name, nc))
            gpu_av[nc] = av_g
        prob.close()
        datas_g.append(X.xensemble(1, q, path_fn=lambda p_, j, nc, npar: gpu_av[nc])[0])
        datas_r.append(X.xensemble(1, q, path_fn=lambda p_, j, nc, npar, _q=q, _a0=a0:
                                   X.run_path(_q, 1, nc, a0=_a0.reshape(1, -1, 1))[1])[0])
    q4 = X.xpreferences(dict(oracle_problem(p), observe=[lambda a, p: a] * 4))
    q4.output = [lambda o, p, _k=k: np.abs(o[_k]) ** 2 for k in range(4)]
    q4.compare = [lambda p: np.abs(_peregrine(p.x, p.t)) ** 2] * 4
    q4._avshape = X._probe_shapes(q4)
    eg = X.xsim_stats(q4, datas_g)[0]
    er = X.xsim_stats(q4, datas_r)[0]
    print("Peregrine 128 sites: error gpu %.7e oracle %.7e; step %.3e / %.3e; difference %.3e / %.3e"
          % (eg[0], er[0], eg[1], er[1], eg[3], er[3]), flush=True)
    if etol is None:
        # FP32: rounding of the single-precision fields (6e-8 relative, grown by the
        # modulational instability over 2 time units) enters the error vector at the
        # 1e-5 level; each entry (total, step, difference) within 5e-5 of the oracle's
        for i in (0, 1, 3):
            assert abs(eg[i] - er[i]) <= 5e-5, (i, eg[i], er[i])
        return
    assert abs(eg[0] - er[0]) <= etol * er[0]
    assert abs(eg[3] - er[3]) <= etol * er[3]


def _kfilter_n(w, p):
    """A filter that is neither real nor symmetric in k (xnoise.m:44)."""
    k = p.kx if p.dimensions == 2 else p.kx - 0.5 * p.ky
    return w * np.exp(-0.5 * (k - 0.3) ** 2) * (1 + 0.2j * k)


def _kfilter_r(v, p):
    k = p.kx if p.dimensions == 2 else p.kx + p.ky
    return v * np.exp(-0.2 * k ** 2) * (0.8 - 0.3j * np.tanh(k))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 5e-5)])
@pytest.mark.parametrize("method", ["RK4", "MP"])
@pytest.mark.parametrize("nx", [64, 256])
def test_nls1d_kspace_noise(prec, tol, method, nx):
    """knoises/krandoms (xnoise.m:39-50, xrandom.m:24-29): filtered k-space
    normals at every k site, inverse transformed inside the kernel once per step
    (coarse steps: mean of the two fine draws) and once for the initial state,
    against the oracle (pinned on the filtered-noise covariances)."""
    p = shrink(nls1d(prec, nx=nx), E=20, e2=2, points=3, steps=2)
    p["method"] = method
    p.update(noises=0, knoises=2, inrandoms=0, krandoms=2, nfilter=_kfilter_n, rfilter=_kfilter_r)
    p["gpu"]["observe"] = ["abs2", "a"]
    _path_case(p, tol)


@pytest.mark.parametrize("prec,tol,n,ini", [("fp64", 1e-10, 64, "phase"), ("fp32", 5e-5, 128, "phase"),
                                            ("fp64", 1e-10, 32, "coherent"), ("fp32", 5e-5, 64, "coherent")])
def test_cgl2d_kspace_noise(prec, tol, n, ini):
    """2D: the k sites of both dimensions in the transforms' frequency order."""
    p = shrink(cgl2d(prec, n=n), E=3, e2=2, points=3, steps=2)
    p.update(noises=0, knoises=2, nfilter=_kfilter_n)
    if ini == "coherent":
        p.update(inrandoms=0, krandoms=2, rfilter=_kfilter_r, urandoms=0)
        p["gpu"] = dict(p["gpu"], initial="coherent", alpha=0.3 + 0.1j)
    p["gpu"]["observe"] = ["abs2_ave", "abs2"] + (["a_ave"] if prec == "fp64" else [])
    _path_case(p, tol)


def test_kspace_noise_one_component():
    """knoises = 1: the drift's noise is s w1 with w1 complex (xift of a filtered real field)."""
    p = shrink(nls1d("fp64", nx=128), E=20, e2=1, points=3, steps=2)
    p.update(noises=0, knoises=1, nfilter=_kfilter_n)
    _path_case(p, 1e-10, nc_list=(1, 2), j=1)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 5e-5)])
@pytest.mark.parametrize("nx,bnd,method", [(101, "NN", "RK4"), (101, "NN", "MP"), (100, "DD", "RK4"),
                                           (77, "DN", "MP"), (45, "ND", "RK4"), (161, "DD", "MP"),
                                           (101, "PP", "RK4")])
def test_nls1d_padded_lattice(prec, tol, nx, bnd, method):
    """1D lattices of any size (the reference's examples use 101, 161, ... points):
    the dense-transform kernel of the next compiled size with the physical
    matrices in its top-left corner (padding sites are neither read nor
    written by the physical ones), Dirichlet ends at the physical last site,
    outputs over the physical sites only; odd periodic lattices use the padded
    DFT matrices.  Boundary values on, against the oracle on the same lattice."""
    p = shrink(nls1d(prec, nx=nx), E=20, e2=2, points=3, steps=2)
    p["method"] = method
    p["boundaries"] = [BOUNDS[bnd]]
    if bnd != "PP":
        p["boundval"] = [[0.2 - 0.1j, 0.15j]]
    p["gpu"]["observe"] = ["abs2", "a", "abs2_ave", "a_ave"]
    _path_case(p, tol)


def _raw_known_answer(p, obs, outputs=None, compare=None, **extra_q):
    """The xsim error vector of a deterministic project (one trajectory) whose
    observables are the reference's own functions of the fields: the GPU path
    keeps the raw fields at every stored point (xpath.m:73-78) and the
    observables, averages and statistics follow xdata/xensemble/xsim."""
    prob = Problem(dict(p, raw=True))
    prob.handle.run_ensemble(1)
    prob.handle.synchronize()
    raw = prob.handle.read_raw()                       # [e2, nchk, points, E, ns]
    prob.close()
    extra = dict(observe=obs, **extra_q)
    if outputs is not None:
        extra["output"] = outputs
    if compare is not None:
        extra["compare"] = compare
    q = X.xpreferences(dict(oracle_problem(p), **extra))
    npts = q.points[0]
    shapes = [X.xdata(np.zeros((1,) + q.space + (1,), complex), o, q).shape for o in range(len(obs))]

    def gpu_path(q_, j, nc, npar):
        av = [np.zeros((sh[0], npts, sh[1]), complex) for sh in shapes]
        for t in range(npts):
            # raw sites x + nx*y -> [field, x, (y), E]
            a = raw[j - 1, nc - 1, t].T.reshape(tuple(reversed(q_.space)) + (-1,))
            a = np.transpose(a, tuple(range(len(q_.space) - 1, -1, -1)) + (len(q_.space),))[None]
            for o in range(len(obs)):
                av[o][:, t, :] = X.xdata(a, o, q_)
        return av

    data = X.xensemble(1, q, path_fn=gpu_path)
    q._avshape = X._probe_shapes(q)
    return X.xsim_stats(q, data)[0]


_NLS_GPU = dict(deriv="gp", g=1j, s=0.0, linear=dict(l0=-0.5j, l2=0.5j), initial="sech", amp=1.0, width=1.0,
                eps=0.0, observe=["a"], precision="fp64")


@pytest.mark.parametrize("row", ["Soliton#4", "Characteristic#8", "SolitonAv#14", "GaussianAv2D#16",
                                 "GaussianKinetic#19", "SolitonDerivN#21"])
def test_batchtest_known_answers_on_gpu(row):
    """Deterministic Batchtest rows run on the GPU path (FP64), tolerance 0.1 %
    (Batchtest.m:67-81): Soliton #4 (35 periodic sites: padded DFT, RK4, no
    checks), Characteristic #8 (first-derivative propagator, shifted sech, a
    compare at one site), SolitonAv #14 (lattice average and integral of Re a,
    RK4 with checks), GaussianAv2D #16 (35 x 35 periodic sites padded to the
    64 x 64 dense kernel, k-space observable, xcompress axes), GaussianKinetic
    #19 (70 x 34 periodic sites padded to the 128 x 64 dense kernel, anisotropic
    Gaussian, k-space integrals with dk), SolitonDerivN #21 (101 sites, Neumann
    DCT-I padded to 128, order-4 extrapolation, Int(|da/dx|^2)).  The
    observables are the reference's own functions of the fields, evaluated on
    the GPU's raw fields."""
    sech = lambda x: 1 / np.cosh(x)
    if row == "Characteristic#8":           # Examples/SPDE_2D/Characteristic.m:12-22
        p = dict(name="Characteristic", dimensions=2, points=[51, 35], noises=0, inrandoms=0,
                 gpu=dict(deriv="zero", linear=dict(l1=-1.0), initial="sech", amp=1.0, width=0.5, center=-2.5,
                          observe=["a"], precision="fp64"))
        err = _raw_known_answer(p, [lambda a, p: a, lambda a, p: a ** 2],
                                compare=[None, lambda p: sech(2 * (p.t - 2.5)) ** 2], axes=[None, [0, 18]])
        expected = 9.654854e-05
    elif row == "GaussianKinetic#19":       # Examples/SPDE_3D/GaussianKinetic.m:17-42
        from test_oracle_batchtest import X0, Y0, gaussiankinetic_observe
        p = dict(name="GaussianKinetic", dimensions=3, ranges=[5, 200, 6], points=[49, 70, 34], noises=0,
                 inrandoms=0, gpu=dict(deriv="zero", linear=dict(l2=[1j, 0]), initial="gauss",
                                       amp=1 / np.sqrt(np.pi * X0 * Y0), width=X0, width_y=Y0, observe=["abs2"],
                                       precision="fp64"))
        ob = gaussiankinetic_observe()
        err = _raw_known_answer(p, ob["observe"], compare=ob["compare"], transforms=ob["transforms"])
        expected = 4.049843e-06
    elif row == "Soliton#4":                # Examples/SPDE_2D/Soliton.m
        p = dict(name="Soliton", dimensions=2, checks=0, order=2, noises=0, inrandoms=0, gpu=dict(_NLS_GPU))
        err = _raw_known_answer(p, [lambda a, p: a], compare=[lambda p: sech(p.x)])
        expected = 5.047219e-03
    elif row == "SolitonAv#14":             # Examples/SPDE_2D/SolitonAv.m
        p = dict(name="SolitonAv", dimensions=2, noises=0, inrandoms=0, gpu=dict(_NLS_GPU))
        err = _raw_known_answer(p, [lambda a, p: X.Ave(np.real(a), p), lambda a, p: X.Int(np.real(a), p)],
                                compare=[lambda p: np.pi / 10 + 0 * p.t, lambda p: np.pi + 0 * p.t])
        expected = 1.410144e-02
    elif row == "GaussianAv2D#16":          # Examples/SPDE_3D/GaussianAv2D.m
        p = dict(name="GaussianAv2D", dimensions=3, noises=0, inrandoms=0,
                 gpu=dict(deriv="zero", linear=dict(l2=[0.05j, 0.05j]), initial="gauss", amp=1.0, width=1.0,
                          observe=["abs2"], precision="fp64"))

        def g2(p):
            sig = 1 + (p.t / 10) ** 2
            return sig ** -1 * np.exp(-(p.x ** 2 + p.y ** 2) / sig)
        ii = lambda a, p: a * np.conj(a)
        err = _raw_known_answer(p, [ii, lambda a, p: X.Ave(ii(a, p), p), lambda a, p: X.Int(ii(a, p), p), ii],
                                compare=[g2, None, None, lambda r: 1], transforms=[0, 0, 0, [0, 1, 1]],
                                axes=[None, None, None, [0, -1, -1]])
        expected = 4.813746e-07
    else:                                   # Examples/SPDE_2D/SolitonDerivN.m
        p = dict(name="SolitonDerivN", dimensions=2, points=[101, 101], ranges=[10.0, 15.0], method="RK4",
                 order=4, noises=0, inrandoms=0, boundaries=[[-1, -1]], gpu=dict(_NLS_GPU))
        err = _raw_known_answer(p, [lambda a, p: a * np.conj(a),
                                    lambda a, p: X.Int(np.abs(X.D1(a, 2, p)) ** 2, p)])
        expected = 2.134464e-06
    print("%s on the GPU: error %.7e (Batchtest %.6e)" % (row, err[0], expected), flush=True)
    assert expected / 1.001 < err[0] < expected * 1.001


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 5e-5)])
@pytest.mark.parametrize("shape,bnd,method", [((35, 35), ("PP", "PP"), "RK4"), ((48, 40), ("NN", "DD"), "MP"),
                                              ((33, 64), ("DN", "PP"), "RK4"), ((64, 21), ("ND", "NN"), "MP"),
                                              ((70, 34), ("PP", "PP"), "RK4"), ((34, 100), ("NN", "DD"), "MP"),
                                              ((128, 41), ("DN", "PP"), "RK4"), ((64, 128), ("ND", "NN"), "MP")])
def test_cgl2d_padded_lattice(prec, tol, shape, bnd, method):
    """2D lattices of sizes without a compiled kernel (odd, unequal) on the 64 x 64,
    128 x 64 and 64 x 128 dense-transform kernels, padding per dimension (and the
    two rectangular kernels at their own sizes): the counter streams and
    outputs keep the physical site indices x + nx*y, boundary values at the
    physical ends; against the oracle on the same lattice."""
    p = shrink(cgl2d(prec, n=32), E=3, e2=2, points=3, steps=2)
    p["points"] = [p["points"][0], shape[0], shape[1]]
    p["method"] = method
    p["boundaries"] = [BOUNDS[b] for b in bnd]
    if bnd != ("PP", "PP"):
        p["boundval"] = [[0.02, -0.01 + 0.005j], [0.05 - 0.02j, 0.03]]
    p["gpu"]["observe"] = ["abs2_ave", "abs2", "a"] + (["a_ave"] if prec == "fp64" else [])
    _path_case(p, tol)


def test_peregrine_known_answer():
    """Examples/SPDE_2D/Peregrine.m itself on the GPU (Batchtest #42: 4.588574e-04,
    tolerance 0.1 %): 161 sites (padded to the 256-site FP64 dense-transform
    kernel), the Peregrine solution's time-dependent values at both ends for
    DD, NN, DN and ND, MP with checks and order-2 extrapolation, FP64; the four
    one-field runs share one xsim statistics pass (the reference's four
    uncoupled field components)."""
    ends = {"DD": [1, 1], "NN": [-1, -1], "DN": [1, -1], "ND": [-1, 1]}
    datas = []
    for name, b in ends.items():
        f = _PEREGRINE_ENDS[name]
        p = dict(name="Peregrine-" + name, dimensions=2, noises=1, ranges=[10.0, 10.0], origins=[-5.0, -5.0],
                 points=[51, 161], method="MP", order=2, steps=20, checks=1, ensembles=[1, 1, 1],
                 boundaries=[b], boundfun=lambda a, c, d, p, _f=f: _f(p.origins[1], p.t),
                 gpu=dict(deriv="gp", g=1j, s=0.0, linear=dict(l2=0.5j), initial="const", c0=0.0,
                          observe=["a"], precision="fp64"))
        q = X.xpreferences(oracle_problem(p))
        q.output = [lambda o, p: np.abs(o[0]) ** 2]
        a0 = _peregrine(q.xc[1], -5.0)
        prob = Problem(p)
        gpu_av = {nc: prob.xpath(nc, npar=1, j=1, a0=a0[None, :]) for nc in (1, 2)}
        prob.close()
        datas.append(X.xensemble(1, q, path_fn=lambda p_, j, nc, npar: gpu_av[nc])[0])
    q4 = X.xpreferences(dict(oracle_problem(p), observe=[lambda a, p: a] * 4))
    q4.output = [lambda o, p, _k=k: np.abs(o[_k]) ** 2 for k in range(4)]
    q4.compare = [lambda p: np.abs(_peregrine(p.x, p.t)) ** 2] * 4
    q4._avshape = X._probe_shapes(q4)
    err = X.xsim_stats(q4, datas)[0]
    print("Peregrine on the GPU: error %.7e (Batchtest 4.588574e-04)" % err[0], flush=True)
    assert 4.588574e-04 / 1.001 < err[0] < 4.588574e-04 * 1.001


@pytest.mark.parametrize("bnd,method", [(("NN", "NN"), "RK4"), (("NN", "PP"), "MP"), (("PP", "NN"), "RK4")])
def test_stream_dense_fp64_128(bnd, method):
    """FP64 128x128 with Neumann ends (config 3's Neumann run in FP64): the
    fields stay in the streaming kernel's workspace and the dense DCT-I (and,
    on mixed lattices, DFT) passes run on row halves and column halves in LDS;
    against the oracle, fine and coarse paths."""
    p = shrink(cgl2d("fp64", n=128), E=2, e2=2, points=3, steps=2)
    p["method"] = method
    p["boundaries"] = [BOUNDS[b] for b in bnd]
    p["gpu"]["observe"] = ["abs2_ave", "abs2", "a_ave", "corr"]
    prob = Problem(p)
    assert prob.info.kernel_kind == 1
    prob.close()
    _path_case(p, 1e-10)


@pytest.mark.parametrize("method", ["RK4", "MP"])
@pytest.mark.parametrize("values", ["boundval", "boundfun"])
def test_stream_dense_fp64_128_boundary_values(method, values):
    """FP64 128x128 with Dirichlet ends and non-zero boundary values on the
    streaming kernel (fields in device memory): Neumann x with derivative
    values, Dirichlet y with field values, constant (p.boundval) or changing
    along the lines and in time (p.boundfun): the patch around every
    propagation, the Neumann correction and the Dirichlet sites at the
    method's stage times, against the oracle, fine and coarse paths."""
    p = shrink(cgl2d("fp64", n=128), E=2, e2=2, points=3, steps=2)
    p["method"] = method
    p["boundaries"] = [[-1, -1], [1, 1]]
    if values == "boundval":
        p["boundval"] = [[0.02, -0.01 + 0.005j], [0.05 - 0.02j, 0.03]]
    else:
        p["boundfun"] = _ends_2d
    p["gpu"]["observe"] = ["abs2_ave", "abs2", "a_ave"]
    prob = Problem(p)
    assert prob.info.kernel_kind == 1
    prob.close()
    _path_case(p, 1e-10)


@pytest.mark.parametrize("method", ["RK4", "MP"])
@pytest.mark.parametrize("randoms", [False, True])
def test_stream_fp64_128_kspace_noise(method, randoms):
    """k-space filtered noises (and initial randoms) on the FP64 128x128
    streaming kernel: the step's x-space noise plane in the workspace, produced
    by the streaming transform with the filtered k-space normals as its middle
    operation; against the oracle's xnoise / xrandom (xift of the filtered
    normals), fine and coarse paths."""
    p = shrink(cgl2d("fp64", n=128), E=2, e2=2, points=3, steps=2)
    p["method"] = method
    p.update(noises=0, knoises=2, nfilter=_kfilter_n)
    if randoms:
        p.update(inrandoms=0, krandoms=2, rfilter=_kfilter_r, urandoms=0)
        p["gpu"] = dict(p["gpu"], initial="coherent", alpha=0.3 + 0.1j)
    p["gpu"]["observe"] = ["abs2_ave", "abs2", "a_ave"]
    prob = Problem(p)
    assert prob.info.kernel_kind == 1
    prob.close()
    _path_case(p, 1e-10)


def test_unsupported_combinations_rejected():
    """Combinations outside the GPU path fail at creation with XSH_EUNSUPPORTED
    (the host then keeps the project on the CPU path), before any launch."""
    from xspde_matlab_amd._lib import XshError
    cases = []
    p = shrink(nls1d("fp64", nx=101), E=4, e2=1, points=2, steps=1)
    p["boundaries"] = [[-1, -1]]
    p["gpu"]["observe"] = ["corr"]                  # periodic FFT of a padded lattice
    cases.append(p)
    p = shrink(cgl2d("fp32", n=64), E=2, e2=1, points=2, steps=1)
    p["boundaries"] = [[-1, -1], [-1, -1]]
    p.update(noises=0, knoises=2, nfilter=_kfilter_n)    # k-space noise on a dense-transform lattice
    cases.append(p)
    for p in cases:
        with pytest.raises(XshError, match="XSH_EUNSUPPORTED"):
            Problem(p)


def test_averages_in_matlab_order():
    """The C ABI's av_out is p.d.av{o} = [lines, points, nx, ny] in MATLAB
    column-major order (xpreferences.m:340): per-site observables time fastest.
    Read through xsh_run_path directly (no host reshaping) and compared with the
    oracle's xdata averages flattened in that order."""
    import ctypes as C
    from xspde_matlab_amd import _lib
    p = shrink(cgl2d("fp64", n=32), E=3, e2=1, points=4, steps=2)
    p["gpu"]["observe"] = ["abs2", "abs2_ave", "a"]
    prob = Problem(p)
    info = prob.info
    av = np.zeros(2 * info.avlen)
    _lib.check(_lib.load().xsh_run_path(prob.handle._h, 1, 1, 1, None, None, av.ctypes.data_as(C.POINTER(C.c_double))))
    prob.close()
    av = av[0::2] + 1j * av[1::2]
    q, _, av_r = _oracle_path(p, 1, 1)
    npts = p["points"][0]
    for o in range(3):
        ref = av_r[o]                                     # [lines, points, sites], sites in C order of (x, y)
        if ref.shape[2] > 1:
            ref = ref.reshape(1, npts, q.space[0], q.space[1])   # [lines, points, nx, ny]
        want = ref.reshape(-1, order="F")
        got = av[info.obs_offset[o]:info.obs_offset[o] + want.size]
        _cmp(got, want, 1e-10, "av[%d] in p.d.av order" % o)
    # time fastest: the second element of the |a|^2 block is site (0, 0) at the second point
    assert np.isclose(av[info.obs_offset[0] + 1], av_r[0][0, 1, 0], rtol=1e-10)


def _oracle_snapshots(q, j, nc):
    """The oracle path's fields at every stored point (xpath.m:73-78)."""
    snaps = []
    fs = list(q.observe)
    q2 = q.copy()
    q2.observe = [lambda a, p, _f=fs[0]: (snaps.append(np.array(a)), _f(a, p))[1]] + fs[1:]
    X.run_path(q2, j, nc)
    return snaps


@pytest.mark.parametrize("case,tol", [("cgl2d-fp64-32", 1e-10), ("nls1d-fp32-64", 5e-5), ("cgl2d-fp64-128", 1e-10)])
def test_raw_fields(case, tol):
    """p.raw: every trajectory's field at every stored point, for every (j, nc)
    path of the ensemble, against the oracle's fields at its observation points
    (register kernels, a partial 1D trajectory group, the streaming kernel)."""
    kind, prec, n = case.split("-")
    n = int(n)
    if kind == "nls1d":
        p = shrink(nls1d(prec, nx=n), E=20, e2=2, points=3, steps=2)
    else:
        p = shrink(cgl2d(prec, n=n), E=3, e2=2, points=3, steps=2)
    p["raw"] = True
    prob = Problem(p)
    prob.handle.run_ensemble(1)
    prob.handle.synchronize()
    raw = prob.handle.read_raw()                       # [e2, nchk, points, E, ns]
    prob.close()
    q = X.xpreferences(oracle_problem(p))
    for j in (1, 2):
        for nc in (1, 2):
            snaps = _oracle_snapshots(q, j, nc)
            assert len(snaps) == p["points"][0]
            for t, a in enumerate(snaps):
                ref = np.moveaxis(a[0], -1, 0)
                ref = ref.reshape(ref.shape[0], -1, order="F")
                got = raw[j - 1, nc - 1, t]
                err = np.abs(got - ref).max() / np.abs(ref).max()
                assert err < tol, (j, nc, t, err)
    print("%s raw fields: %d paths x %d points match" % (case, 4, p["points"][0]))


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("shape", [(), (64,), (128, 128)])
def test_divergence_flag(prec, shape):
    """The kernels' non-finite check: a drift g |a|^2 a with g = 50 blows up
    within the path (the flag is set and xsim warns); g = -1 stays finite."""
    import warnings as W
    from xspde_matlab_amd.xsim import DivergenceWarning
    nd = len(shape) + 1
    p = dict(name="Blowup", dimensions=nd, points=[5] + list(shape), ranges=[2.0] + [10.0] * len(shape),
             steps=[2] + [1] * len(shape), checks=0, noises=0, method="RK4", ensembles=[2, 1, 1],
             gpu=dict(deriv="gp", g=50.0, u=0.0, s=0.0, initial="const", c0=1.0, observe=["abs2_ave"],
                      precision=prec))
    prob = Problem(p)
    with W.catch_warnings(record=True) as rec:
        W.simplefilter("always")
        prob.xensemble(1)
    assert prob.handle.nonfinite()
    assert any(issubclass(r.category, DivergenceWarning) for r in rec)
    prob.close()
    p["gpu"]["g"] = -1.0
    prob = Problem(p)
    prob.xensemble(1)
    assert not prob.handle.nonfinite()
    prob.close()
