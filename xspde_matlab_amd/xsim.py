"""Host mirror of xSPDE's ``xsim``/``xensemble``/``xpath`` for the GPU path.

A problem is written like an xSPDE project (a ``p`` dictionary with the
reference's field names, ``doc/source/api.rst``) plus a ``gpu`` entry that
names the built-in forms the GPU integrates::

    p = dict(dimensions=3, points=[101, 128, 128], ranges=[4, 20, 20],
             steps=[20, 1, 1], method='RK4', noises=2, checks=1,
             ensembles=[64, 16, 1],
             gpu=dict(deriv='gp', g=-(1 + 0.5j), u=0, s=0.05 / 2 ** 0.5,
                      linear=dict(l0=1, l2=1 + 1j), initial='phase', amp=0.1,
                      observe=['abs2_ave', 'corr'], precision='fp32'))
    error, data, p, raw = xsim(p)

Names follow the reference: ``xpath`` is the p.path contract
(Simulator/xpath.m:1), ``xensemble`` the sub-ensemble/check loops with
D, D^2 accumulation (Simulator/xensemble.m:12-108), ``xsim`` the parallel
ensemble loop and the final error statistics (Simulator/xsim.m:28-188).
A drift given as a Python callable instead of a ``gpu`` form is rejected:
in the MATLAB host such projects stay on the reference's CPU ``xpath``.
A project with several field components (``fields > 1``) runs each
component as its own handle (``FieldsProblem``): the built-in forms do not
couple components.
"""
import math
import time
import warnings

import numpy as np

from . import _lib


class GpuFormError(ValueError):
    """The project cannot run on the GPU path (not a built-in form)."""


class DivergenceWarning(RuntimeWarning):
    """The kernels saw a NaN or infinite field value at a stored time point."""


def _vec(x, n, default):
    if x is None:
        x = default
    if np.isscalar(x):
        x = [x]
    x = list(x)
    while len(x) < n:
        x.append(default if np.isscalar(default) else x[-1])
    return x[:n]


def preferences(p):
    """Defaults and derived sizes of the fields the path reads (xpreferences.m:42-175)."""
    q = dict(p)
    if "gpu" not in q:
        raise GpuFormError("no 'gpu' forms: the project stays on the CPU xpath")
    for k in ("deriv", "linear", "initial", "observe"):
        if callable(q.get(k)) or (k == "observe" and q.get(k) is not None):
            raise GpuFormError("user function handle '%s': not a built-in GPU form" % k)
    nd = max(1, int(q.get("dimensions", 1)))
    q["dimensions"] = nd
    fields = int(q.get("fields", 1))
    if fields < 1:
        raise GpuFormError("fields must be at least 1")
    q["fields"] = fields
    # cells the project names (an empty one counts as not named, as [] in MATLAB)
    given = {k: q.get(k) is not None for k in ("knoises", "unoises", "urandoms", "krandoms")}
    q["knoises"] = int(q.get("knoises") or 0)
    q["krandoms"] = int(q.get("krandoms") or 0)
    if q.get("noises") is None and not given["knoises"] and not given["unoises"]:
        q["noises"] = fields                                         # xpreferences.m:65-70
    q["noises"] = int(q.get("noises") or 0)
    if q.get("inrandoms") is None and not given["urandoms"] and not given["krandoms"]:
        q["inrandoms"] = q["noises"]                                 # xpreferences.m:71-76
    q["inrandoms"] = int(q.get("inrandoms") or 0)
    q["urandoms"] = int(q.get("urandoms") or 0)
    q["ensembles"] = [int(e) for e in _vec(q.get("ensembles"), 3, 1)]
    q["ranges"] = [float(r) for r in _vec(q.get("ranges"), nd, 10.0)]
    steps = [int(s) for s in _vec(q.get("steps"), nd, 1)]
    q["steps"] = steps
    pts = q.get("points")
    pts = [51] + [35] * (nd - 1) if pts is None else _vec(pts, nd, 35)             # xpreferences.m:56
    pts = [int(pts[0])] + [1 + (int(pts[d]) - 1) * steps[d] for d in range(1, nd)]   # :137
    q["points"] = pts
    og = q.get("origins")
    og = [0.0] if og is None else [float(o) for o in np.ravel(og)]
    q["origins"] = (og + [-r / 2 for r in q["ranges"][len(og):]])[:nd]        # xpreferences.m:83
    ch = q.get("checks")
    ch = [1] + [0] * (nd - 1) if ch is None else _vec(ch, nd, 0)
    if any(ch[1:]):
        raise GpuFormError("spatial checks are not on the GPU path")
    q["checks"] = [int(c) for c in ch]
    meth = q.get("method")
    if meth is None:
        meth = "RK4" if np.prod(q["ensembles"]) == 1 else "MP"       # xpreferences.m:101-105
    if not isinstance(meth, str):
        meth = getattr(meth, "__name__", str(meth))
    if meth not in ("RK4", "MP"):
        raise GpuFormError("method %s is not on the GPU path" % meth)
    q["method"] = meth
    q["iterations"] = int(q.get("iterations", 4))
    q["order"] = int(q.get("order", 0))
    q["seed"] = int(q.get("seed", 0))
    q["dt"] = q["ranges"][0] / max(1, pts[0] - 1)                   # :174
    q["rawdata"] = int(bool(q.get("rawdata", q.get("raw", 0))))      # :85
    q["file"] = q.get("file") or " "                                 # :90
    q["version"] = q.get("version") or "xSPDE4.21"                   # :19
    return q


def coupled(q):
    """Several components coupled by the matrix drift form (one descriptor, SDEs)."""
    return int(q.get("fields", 1)) > 1 and q.get("gpu", {}).get("deriv") == "matrix"


def make_desc(q):
    if q.get("fields", 1) != 1 and not coupled(q):
        raise GpuFormError("one descriptor integrates one field component (FieldsProblem splits a project)")
    g = q["gpu"]
    d = _lib.Desc()
    d.precision = _lib.FP64 if str(g.get("precision", "fp32")).lower() in ("fp64", "double") else _lib.FP32
    nd = q["dimensions"]
    d.dimensions = nd
    d.points = q["points"][0]
    d.nx = q["points"][1] if nd > 1 else 1
    d.ny = q["points"][2] if nd > 2 else 1
    d.steps = q["steps"][0]
    d.checks = q["checks"][0]
    d.method = _lib.RK4 if q["method"] == "RK4" else _lib.MP
    d.iterations = q["iterations"]
    for i in range(3):
        d.ensembles[i] = q["ensembles"][i]
    d.seed = q["seed"]
    d.noises, d.inrandoms, d.urandoms = q["noises"], q["inrandoms"], q["urandoms"]
    bnd = q.get("boundaries") or []
    for k in range(2):
        if k < len(bnd) and bnd[k] is not None:
            d.boundaries[k][0], d.boundaries[k][1] = int(bnd[k][0]), int(bnd[k][1])
    for i in range(min(3, nd)):
        d.ranges[i] = q["ranges"][i]
        d.origins[i] = q["origins"][i]
    kind = g.get("deriv", "zero")
    if kind not in _lib.DERIV:
        raise GpuFormError("unknown drift form %r" % (kind,))
    d.deriv = _lib.DERIV[kind]
    if kind == "kubo":
        d.deriv_par[0] = float(g.get("c", 1.0))
    elif kind == "gp":
        gg, u = complex(g.get("g", 0)), complex(g.get("u", 0))
        d.deriv_par[0], d.deriv_par[1], d.deriv_par[2], d.deriv_par[3] = gg.real, gg.imag, u.real, u.imag
        d.deriv_par[4] = float(g.get("s", 0.0))
    lin = g.get("linear")
    if kind == "matrix":
        # coupled SDE components: da_i/dt = sum_j M_ij a_j + g_i |a_i|^2 a_i + sum_k S_ik w_k
        F = int(q.get("fields", 1))
        if F < 2 or F > _lib.XSH_MAX_FIELDS:
            raise GpuFormError("the matrix drift form couples 2..%d components" % _lib.XSH_MAX_FIELDS)
        d.fields = F
        M = np.asarray(g.get("M", np.zeros((F, F))), complex).reshape(F, F)
        gn = np.broadcast_to(np.asarray(g.get("g", 0), complex), (F,))
        S = np.asarray(g.get("S", np.zeros((F, max(1, q["noises"])))), float).reshape(F, -1)
        if S.shape[1] > 4 or q["noises"] > 4:
            raise GpuFormError("the matrix drift form takes at most 4 noise components")
        l0 = np.broadcast_to(np.asarray((lin or {}).get("l0", 0), complex), (F,))
        c0 = np.broadcast_to(np.asarray(g.get("c0", 1.0), complex), (F,))
        for i in range(F):
            for j in range(F):
                d.fmat[i][j][0], d.fmat[i][j][1] = M[i, j].real, M[i, j].imag
            d.fnl[i][0], d.fnl[i][1] = gn[i].real, gn[i].imag
            for k in range(S.shape[1]):
                d.fnoise[i][k] = S[i, k]
            d.fl0[i][0], d.fl0[i][1] = l0[i].real, l0[i].imag
            d.fc0[i][0], d.fc0[i][1] = c0[i].real, c0[i].imag
        d.has_linear = 1 if lin is not None else 0
        of = [int(f) for f in (g.get("ofield") or [1] * len(g.get("observe", ["a"])))]
        for o, f in enumerate(of[:_lib.XSH_MAX_OBS]):
            d.obs_field[o] = f - 1
        lin = None
    if lin is not None:
        d.has_linear = 1
        l0 = complex(lin.get("l0", 0))
        d.lin_l0[0], d.lin_l0[1] = l0.real, l0.imag
        c1 = np.broadcast_to(np.asarray(lin.get("l1", 0), complex), (2,))
        c2 = np.broadcast_to(np.asarray(lin.get("l2", 0), complex), (2,))
        for k in range(2):
            d.lin_l1[k][0], d.lin_l1[k][1] = c1[k].real, c1[k].imag
            d.lin_l2[k][0], d.lin_l2[k][1] = c2[k].real, c2[k].imag
    ini = g.get("initial", "const")
    if ini not in _lib.INITIAL:
        raise GpuFormError("unknown initial form %r" % (ini,))
    d.initial = _lib.INITIAL[ini]
    if ini == "const":
        c = complex(np.ravel(g.get("c0", 1.0))[0])
        d.init_par[0], d.init_par[1] = c.real, c.imag
    elif ini == "sech":
        d.init_par[0], d.init_par[1], d.init_par[2] = g.get("amp", 1.0), g.get("width", 1.0), g.get("eps", 0.0)
        d.init_par[3] = g.get("center", 0.0)
    elif ini == "gauss":
        d.init_par[0], d.init_par[1] = g.get("amp", 1.0), g.get("width", 1.0)
        d.init_par[2] = g.get("width_y", 0.0)
    elif ini == "phase":
        d.init_par[0] = g.get("amp", 1.0)
    elif ini == "coherent":
        a = complex(g.get("alpha", 1.0))
        d.init_par[0], d.init_par[1] = a.real, a.imag
    obs = list(g.get("observe", ["a"]))
    if not 1 <= len(obs) <= _lib.XSH_MAX_OBS:
        raise GpuFormError("1..%d observables" % _lib.XSH_MAX_OBS)
    d.nobserve = len(obs)
    for o, name in enumerate(obs):
        if name not in _lib.OBSERVE:
            raise GpuFormError("unknown observe form %r" % (name,))
        d.observe[o] = _lib.OBSERVE[name]
    # p.transforms{o}(1): observable o of the time-domain spectrum (xpath.m:90-106)
    tr = q.get("transforms")
    if tr is not None:
        tr = list(tr) if isinstance(tr, (list, tuple)) else [tr]
        for o in range(len(obs)):
            t = tr[o] if o < len(tr) else 0
            d.tspec[o] = 1 if (np.ravel(t)[0] if np.size(t) else 0) else 0
    bv = q.get("boundval") or []
    for k in range(min(len(bv), 2)):
        if bv[k] is None:
            continue
        for side in range(2):
            v = complex(bv[k][side])
            d.boundval[k][side][0], d.boundval[k][side][1] = v.real, v.imag
    tab = boundary_table(q)
    if tab is not None:
        d.set_boundary_table(tab)
    d.set_filters(filter_table(q, "nfilter", q["knoises"]), filter_table(q, "rfilter", q["krandoms"]))
    d.seq = int(g.get("seq", 0))
    d.inst = int(g.get("inst", 0))
    d.traj_out = 1 if g.get("traj_out") else 0
    d.raw = 1 if q.get("rawdata", q.get("raw")) else 0
    modes = g.get("modes")
    if modes:
        if isinstance(modes, int):
            modes = lowest_modes(d.nx, d.ny, modes)
        if len(modes) > _lib.XSH_MAX_MODES:
            raise GpuFormError("at most %d noise modes" % _lib.XSH_MAX_MODES)
        d.nmodes = len(modes)
        for m, (kx, ky) in enumerate(modes):
            d.modes[m][0], d.modes[m][1] = int(kx), int(ky)
    return d


class Params(dict):
    """The project's parameters with attribute access (the reference's struct p),
    as handed to p.boundfun."""

    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError as e:
            raise AttributeError(k) from e

    def __setattr__(self, k, v):
        self[k] = v


def _boundary_line(v, space, k):
    """One boundary value of space dimension k (0: x, 1: y) as the vector of its
    values along the other dimension: a scalar, such a vector, or an array of the
    lattice's space shape (values taken at the first site along k)."""
    v = np.asarray(v, dtype=complex)
    nother = space[1 - k] if len(space) == 2 else 1
    if v.size == 1:
        return np.full(nother, v.reshape(-1)[0])
    if v.size == int(np.prod(space)):
        v = np.take(v.reshape(space), 0, axis=k)
    if v.size != nother:
        raise GpuFormError("boundary value of size %d does not fit dimension %d" % (v.size, k + 2))
    return v.reshape(-1)


def boundary_table(q):
    """p.boundfun tabulated for the kernels (xpatch.m:37 and xsetbound.m:39 call it
    at the method's p.t): values [level][dim][lower, upper][b] at the times
    origins(1) + m dtr/2, m = 0 .. 2 n (1 + checks), with dtr the fine step and
    n the method calls of one path.  One level when the values do not change in
    time.  boundfun(a, c, d, p) is called with a = None: values that depend on
    the fields cannot be tabulated, and such projects stay on the CPU path."""
    f = q.get("boundfun")
    if f is None:
        return None
    if not callable(f):
        raise GpuFormError("boundfun must be a function of (a, c, d, p)")
    nd = q["dimensions"]
    bnd = q.get("boundaries") or []
    dims = [k for k in range(nd - 1) if k < len(bnd) and bnd[k] is not None and any(int(b) for b in bnd[k])]
    if not dims:
        return None
    space = tuple(q["points"][1:nd])
    nb = max(space) if nd == 3 else 1
    npts, steps, checks = q["points"][0], q["steps"][0], q["checks"][0]
    calls = npts * steps - 1 if npts > 1 else steps
    levels = 2 * calls * (1 + checks) + 1
    dtr = q["dt"] / (steps * (1 + checks))
    pp = Params(q)
    dx = [q["ranges"][d] / max(1, q["points"][d] - 1) for d in range(nd)]
    names = ["x", "y"]
    for k in range(nd - 1):
        shape = [1] * (nd + 1)
        shape[k + 1] = space[k]
        grid = q["origins"][k + 1] + np.arange(space[k]) * dx[k + 1]
        pp[names[k]] = np.broadcast_to(grid.reshape(shape), (1,) + space + (1,))
    tab = np.zeros((levels, 2, 2, nb), complex)
    for m in range(levels):
        pp.t = q["origins"][0] + m * dtr / 2
        for k in dims:
            try:
                b = f(None, 1, k + 2, pp)
            except (TypeError, AttributeError, IndexError) as e:
                raise GpuFormError("boundfun depends on the fields (%s): not tabulated for the GPU" % e)
            for side in range(2):
                line = _boundary_line(b[side], space, k)
                tab[m, k, side, :line.size] = line
    if np.array_equal(tab, np.broadcast_to(tab[:1], tab.shape)):
        tab = tab[:1]
    return tab


def filter_table(q, name, count):
    """The k-space filter p.nfilter / p.rfilter (xnoise.m:44, xrandom.m:26) as its
    multiplier at every k site: [count][nx*ny] complex, site kx + nx*ky in FFT
    order, with the k grids p.kx, p.ky in that order (spdpreferences.m:61-62).
    The default filter is the identity (xpreferences.m:419-425).  A filter that
    is not a site-wise multiplication has no such table: GpuFormError."""
    if not count:
        return None
    nd = q["dimensions"]
    space = tuple(q["points"][1:nd])
    shape = (count,) + space + (1,)
    f = q.get(name)
    if f is None:
        return np.ones((count, int(np.prod(space))), complex)
    pp = Params(q)
    names = ["x", "y"]
    for k in range(nd - 1):
        n = space[k]
        dx = q["ranges"][k + 1] / max(1, n - 1)
        kc = np.fft.ifftshift((np.arange(n) - n // 2) * 2 * np.pi / (n * dx))
        sh = [1] * (nd + 1)
        sh[k + 1] = n
        pp["k" + names[k]] = kc.reshape(sh)
        pp[names[k]] = (q["origins"][k + 1] + np.arange(n) * dx).reshape(sh)
    F = np.broadcast_to(np.asarray(f(np.ones(shape), pp), dtype=complex), shape)
    z = np.random.default_rng(12345).standard_normal(shape)
    if not np.allclose(np.asarray(f(z, pp), dtype=complex), F * z, rtol=1e-12, atol=1e-300):
        raise GpuFormError("%s is not a multiplication by a function of k: no GPU table" % name)
    # [count, x, (y), 1] -> [count, site = kx + nx*ky]
    return np.ascontiguousarray(np.transpose(F[..., 0], (0,) + tuple(range(nd - 1, 0, -1)))).reshape(count, -1)


def lowest_modes(nx, ny, count):
    """The `count` lattice Fourier modes of smallest |k| (ties by angle), as FFT indices."""
    fx = np.arange(nx) - nx // 2
    fy = np.arange(ny) - ny // 2
    cand = [(int(a * a + b * b), float(np.arctan2(b, a)), int(a), int(b)) for a in fx for b in fy]
    cand.sort()
    return [(a % nx, b % ny) for _, _, a, b in cand[:count]]


def split_averages(q, info, av):
    """Flat per-path averages -> av[o] of shape [1, points, sites_o] (xpath.m:16,81).
    Each observable's block is p.d.av{o} = [lines, points, space] in MATLAB
    column-major order (xpreferences.m:340): time fastest, then x, then y."""
    out = []
    npts = q["points"][0]
    for o in range(len(q["gpu"].get("observe", ["a"]))):
        off, ns = info.obs_offset[o], info.obs_sites[o]
        out.append(np.reshape(av[off:off + npts * ns], (1, npts, ns), order="F"))
    return out


def space_dims(q):
    """Lattice sizes of the space dimensions, x first (p.points(2:end))."""
    return [int(n) for n in q["points"][1:q["dimensions"]]]


def _unflatten_sites(x, axis, dims):
    """Split axis (sites, x fastest) of x into the space dimensions."""
    if len(dims) < 2:
        return x
    k = len(dims)
    x = x.reshape(x.shape[:axis] + tuple(reversed(dims)) + x.shape[axis + 1:])   # C order: x last
    perm = list(range(axis)) + list(reversed(range(axis, axis + k))) + list(range(axis + k, x.ndim))
    return np.transpose(x, perm)


def matlab_data(q, zero, outputs, outd):
    """Per output, the [ndata, 3] statistics as the array xsim.m:150-157 stores:
    the output's own shape with space dimensions unflattened when it is a
    per-site observable ([lines, points, space..., 3]), the mean, step error and
    sampling error along the last axis.  ``zero``: the av list of a zero path,
    for the outputs' shapes."""
    dims = space_dims(q)
    ns = int(np.prod(dims)) if dims else 1
    res = []
    for o, out in enumerate(outputs):
        shape = np.asarray(out(zero, q)).shape
        arr = np.asarray(outd[o]).reshape(tuple(shape) + (3,))
        if len(shape) == 3 and shape[2] == ns and ns > 1:
            arr = _unflatten_sites(arr, 2, dims)
        res.append(arr)
    return res


def matlab_raw(q, r):
    """One path's raw fields from read_raw ([points, E, sites]) as p.d.raw:
    [1, points, space..., E] (xpreferences.m:148, xpath.m:73-78)."""
    dims = space_dims(q)
    x = np.transpose(r, (0, 2, 1))                                   # [points, sites, E]
    x = _unflatten_sites(x, 1, dims) if len(dims) > 1 else x
    if not dims:
        x = x[:, 0, :]
    return x[None]


class Problem:
    """A GPU handle plus the host-side bookkeeping of one xSPDE project."""

    def __init__(self, p):
        self.q = preferences(p)
        if self.q["fields"] != 1 and not coupled(self.q):
            raise GpuFormError("Problem integrates one field component (or coupled SDE components); "
                               "FieldsProblem several uncoupled ones")
        self.handle = _lib.Handle(make_desc(self.q))
        self.info = self.handle.info
        outs = self.q.get("output")
        nobs = len(self.q["gpu"].get("observe", ["a"]))
        if outs is None:
            outs = [None] * nobs
        self.outputs = [o if o is not None else (lambda av, p, _n=n: av[_n]) for n, o in enumerate(outs)]

    def xpath(self, nc, npar=1, j=1, a0=None, want_final=False):
        """p.path for one (j, nc): returns av (and the final fields)."""
        r = self.handle.run_path(npar, j, nc, a0=a0, want_final=want_final)
        if want_final:
            return split_averages(self.q, self.info, r[0]), r[1]
        return split_averages(self.q, self.info, r)

    def xensemble(self, npar, want_raw=False):
        """All sub-ensembles and checks of one parallel ensemble (xensemble.m:25-105).
        With want_raw (needs p.rawdata) also returns raw[nc][j], each path's
        fields as p.d.raw."""
        self.handle.run_ensemble(npar)
        self.handle.synchronize()
        if self.handle.nonfinite():
            warnings.warn("%s: a trajectory's field was not finite at a stored time point (parallel "
                          "ensemble %d): its averages hold NaN or Inf" % (self.q.get("name", "project"), npar),
                          DivergenceWarning, stacklevel=2)
        data = ensemble_moments(self.q, self.info, self.outputs, self.handle.read_averages())
        if not want_raw:
            return data
        r = self.handle.read_raw()
        raw = [[matlab_raw(self.q, r[j, nc]) for j in range(r.shape[0])] for nc in range(r.shape[1])]
        return data, raw

    def zero_av(self):
        return split_averages(self.q, self.info, np.zeros(self.info.avlen, dtype=complex))

    def close(self):
        self.handle.close()


def _field_rows(v, F, pair_len=2):
    """One entry per field: v itself when it is a single [lower, upper] pair (or
    None), else v's rows (the reference's fields x 2 p.boundaries{d} and
    p.boundval{d}, spdpreferences.m:17,26-66)."""
    if v is None:
        return [None] * F
    rows = list(v)
    if len(rows) == F and F > 1 and all(isinstance(r, (list, tuple, np.ndarray)) and np.size(r) == pair_len
                                        for r in rows):
        return [list(r) for r in rows]
    return [v] * F


def field_problems(p, q):
    """A project with p.fields = F components split into F one-field projects
    (the built-in drift forms act on each component by itself, so the
    components do not couple).  Field k takes row k of p.boundaries{d} and
    p.boundval{d}, the k-th (lower, upper) pair of a p.boundfun that returns one
    per field, its noise and random components (p.noises / F of them), the
    overrides p.gpu.fields[k] of the gpu forms, and the observables o with
    p.gpu.ofield[o] = k + 1 (default 1).  Its streams are those of instance k.
    p is the project as written, q its preferences."""
    F = q["fields"]
    g = dict(q["gpu"])
    per = list(g.pop("fields", None) or [{}] * F)
    if len(per) != F:
        raise GpuFormError("gpu.fields needs one entry per field (%d)" % F)
    obs = list(g.get("observe", ["a"]))
    ofield = [int(f) for f in (g.pop("ofield", None) or [1] * len(obs))]
    if len(ofield) != len(obs) or any(not 1 <= f <= F for f in ofield):
        raise GpuFormError("gpu.ofield: one field index 1..%d per observable" % F)
    tr = q.get("transforms")
    tr = (list(tr) if isinstance(tr, (list, tuple)) else [tr]) if tr is not None else None
    nd = q["dimensions"]
    bnd = list(q.get("boundaries") or [])
    bval = list(q.get("boundval") or [])
    brows = [_field_rows(bnd[d] if d < len(bnd) else None, F) for d in range(nd - 1)]
    vrows = [_field_rows(bval[d] if d < len(bval) else None, F) for d in range(nd - 1)]
    bf = q.get("boundfun")
    subs, omap = [], []
    for k in range(F):
        sk = dict(p)
        sk["fields"] = 1
        gk = dict(g)
        gk.update(per[k])
        # each component's share of the noise and random cells; cells its forms do
        # not read may be any number (Peregrine.m declares noises = 1 for 4 fields)
        used = {"noises": gk.get("deriv") == "kubo" or (gk.get("deriv") == "gp" and gk.get("s", 0) != 0),
                "knoises": gk.get("deriv") == "kubo" or (gk.get("deriv") == "gp" and gk.get("s", 0) != 0),
                "inrandoms": gk.get("initial") == "coherent" or (gk.get("initial") == "sech" and gk.get("eps", 0)),
                "krandoms": gk.get("initial") == "coherent" or (gk.get("initial") == "sech" and gk.get("eps", 0)),
                "urandoms": gk.get("initial") == "phase"}
        for name in ("noises", "inrandoms", "urandoms", "knoises", "krandoms"):
            if q[name] % F and used[name]:
                raise GpuFormError("%s = %d is not a whole number of components per field (%d fields)"
                                   % (name, q[name], F))
            sk[name] = q[name] // F if used[name] else 0
        mine = [o for o in range(len(obs)) if ofield[o] == k + 1]
        gk["observe"] = [obs[o] for o in mine] or ["a_ave"]
        gk["inst"] = int(g.get("inst", 0)) * F + k
        sk["gpu"] = gk
        if tr is not None:
            sk["transforms"] = [tr[o] if o < len(tr) else 0 for o in mine] or [0]
        sk["boundaries"] = [brows[d][k] for d in range(nd - 1)] if bnd else q.get("boundaries")
        sk["boundval"] = [vrows[d][k] for d in range(nd - 1)] if bval else q.get("boundval")
        if bf is not None:
            def fk(a, c, d, p, _f=bf, _k=k):
                b = _f(a, c, d, p)
                if len(b) == F and F > 1 and isinstance(b[0], (list, tuple)):
                    return b[_k]
                return b
            sk["boundfun"] = fk
        subs.append(sk)
        omap.append(mine)
    where = {}
    for k, mine in enumerate(omap):
        for i, o in enumerate(mine):
            where[o] = (k, i)
    return subs, [where[o] for o in range(len(obs))]


class FieldsProblem:
    """A project with several field components (p.fields > 1,
    xpreferences.m:44-50): one GPU handle per component, launched together on
    their own streams, the averages reassembled in the project's observable
    order.  Same interface as Problem."""

    def __init__(self, p):
        self.q = preferences(p)
        subs, self.where = field_problems(p, self.q)
        self.parts = []
        try:
            for sk in subs:
                self.parts.append(Problem(sk))
        except Exception:
            self.close()
            raise
        self.info = self.parts[0].info
        outs = self.q.get("output")
        nobs = len(self.where)
        if outs is None:
            outs = [None] * nobs
        self.outputs = [o if o is not None else (lambda av, p, _n=n: av[_n]) for n, o in enumerate(outs)]

    def _assemble(self, per_field):
        return [per_field[k][i] for k, i in self.where]

    def xpath(self, nc, npar=1, j=1, a0=None, want_final=False):
        """p.path for one (j, nc): av in the project's observable order; with
        want_final also the final fields [F, E, sites]. a0: [F, E, sites] or None."""
        res = [part.xpath(nc, npar, j, a0=None if a0 is None else a0[k], want_final=want_final)
               for k, part in enumerate(self.parts)]
        if want_final:
            E, ns = self.q["ensembles"][0], self.info.nsites
            return self._assemble([r[0] for r in res]), np.stack([r[1].reshape(E, ns) for r in res])
        return self._assemble(res)

    def xensemble(self, npar, want_raw=False):
        for part in self.parts:
            part.handle.run_ensemble(npar)
        for part in self.parts:
            part.handle.synchronize()
        for k, part in enumerate(self.parts):
            if part.handle.nonfinite():
                warnings.warn("%s: field %d was not finite at a stored time point (parallel ensemble %d): "
                              "its averages hold NaN or Inf" % (self.q.get("name", "project"), k + 1, npar),
                              DivergenceWarning, stacklevel=2)
        avs = [part.handle.read_averages() for part in self.parts]
        split = lambda j, nc: self._assemble([split_averages(part.q, part.info, av[j, nc])
                                              for part, av in zip(self.parts, avs)])
        data = ensemble_moments(self.q, self.info, self.outputs, None, split=split)
        if not want_raw:
            return data
        rs = [part.handle.read_raw() for part in self.parts]
        raw = [[np.concatenate([matlab_raw(part.q, r[j, nc]) for part, r in zip(self.parts, rs)], axis=0)
                for j in range(rs[0].shape[0])] for nc in range(rs[0].shape[1])]
        return data, raw

    def zero_av(self):
        return self._assemble([part.zero_av() for part in self.parts])

    def close(self):
        for part in self.parts:
            part.close()
        self.parts = []


def make_problem(p):
    """Problem for one field component or coupled SDE components, FieldsProblem
    for several uncoupled ones."""
    if int(p.get("fields", 1)) > 1 and not coupled(p):
        return FieldsProblem(p)
    return Problem(p)


def ensemble_moments(q, info, outputs, av_all, split=None):
    """D and D^2 sums of one parallel ensemble from the averages of its paths
    (xensemble.m:84-105): av_all[j, nc] is the flat average of path (j, nc);
    each output's D is extrapolated by (D_fine - D_coarse) / (2^order - 1) when
    checks and order > 0; the sums carry the weight 1 / (ensembles(2) ensembles(3)).
    Returns data[o] of shape [ndata_o, max(3, nchk + 1)]: sum D per check, then
    sum D_fine^2.  ``split(j, nc)``, when given, returns path (j, nc)'s av list
    instead (several field components, each from its own handle)."""
    e2, e3 = q["ensembles"][1], q["ensembles"][2]
    nchk = 1 + q["checks"][0]
    w = 1.0 / (e2 * e3)
    m1 = max(3, nchk + 1)
    data = [None] * len(outputs)
    for j in range(e2):
        D = {}
        for nc in range(nchk):
            av = split(j, nc) if split is not None else split_averages(q, info, av_all[j, nc])
            for o, out in enumerate(outputs):
                D[(nc + 1, o)] = np.real(np.asarray(out(av, q))).reshape(-1)
        for o in range(len(outputs)):
            if q["checks"][0] and q["order"] > 0:                        # xensemble.m:93-97
                dl = (D[(1, o)] - D[(2, o)]) / (2.0 ** q["order"] - 1.0)
                D[(2, o)] = D[(1, o)]
                D[(1, o)] = D[(1, o)] + dl
            if data[o] is None:
                data[o] = np.zeros((D[(1, o)].size, m1))
            for n1 in range(1, nchk + 1):
                data[o][:, n1 - 1] += D[(n1, o)] * w
            data[o][:, m1 - 1] += D[(1, o)] ** 2 * w
    return data


def statistics(q, data, shapes=None):
    """Step and sampling errors of xsim.m:94-108, the comparisons with p.compare
    of :110-143 (when the outputs' shapes are given) and the RMS summary of
    :144-188: returns (error vector, per-output [ndata, 3] arrays of mean, step
    error, sampling error)."""
    er, outd = statistics_rows(q, data, shapes)
    return error_vector(q, er), outd


def compare_grid(q):
    """p.t, p.x, p.y of xgrid.m:49-56 for a compare function: the time points
    along axis 1 and the lattice coordinates along axes 2 and 3 (x space);
    also as p.r, indexed as the reference's p.r{d} (index 0 unused)."""
    nd = q["dimensions"]
    pp = Params(q)
    npts = q["points"][0]
    sh = [1] * (nd + 1)
    sh[1] = npts
    pp["t"] = (q["origins"][0] + np.arange(npts) * q["dt"]).reshape(sh)
    for k, name in enumerate(("x", "y")[:nd - 1]):
        n = q["points"][k + 1]
        s2 = [1] * (nd + 1)
        s2[k + 2] = n
        pp[name] = (q["origins"][k + 1] + np.arange(n) * (q["ranges"][k + 1] / max(1, n - 1))).reshape(s2)
    pp["r"] = [None, pp["t"]] + [pp[name] for name in ("x", "y")[:nd - 1]]   # p.r{1} = t, p.r{d}
    return pp


def _compare_values(q, fn, shape):
    """p.compare{o}(p) at the output's data points, in the order of its rows
    (the output's [lines, points, sites] flattened, sites x fastest)."""
    dims = space_dims(q)
    ns = int(np.prod(dims)) if dims else 1
    c = np.asarray(fn(compare_grid(q)), dtype=float)
    lines, npts = shape[0], shape[1]
    if len(shape) == 3 and shape[2] == ns and ns > 1:       # per-site output
        c = np.broadcast_to(c, (lines, npts) + tuple(dims))
        if len(dims) == 2:
            c = np.transpose(c, (0, 1, 3, 2))               # sites: x fastest
        return c.reshape(-1)
    c = np.broadcast_to(c, (lines, npts) + (1,) * len(dims))
    return c.reshape(-1)


def statistics_rows(q, data, shapes=None):
    """Per-output RMS rows of xsim.m:94-143 (with the comparisons when shapes is
    given and the project has p.compare) and the [ndata, 3] arrays."""
    ne = q["ensembles"][1] * q["ensembles"][2]
    nchk = 1 + q["checks"][0]
    sx = max(3, 1 + nchk)
    tol = float(q.get("tol", 1e-20))
    cutoff = float(q.get("cutoff", -1e20))
    compare = list(q.get("compare") or []) if shapes is not None else []
    er = np.zeros((len(data), max(6, nchk + 4)))
    outd = []
    for o, d in enumerate(data):
        d = np.array(d, dtype=float)
        mx = np.max(np.abs(d[:, 0])) ** 2
        if not (q.get("relerr", 1) and mx > tol):
            mx = 1.0
        if ne > 1:
            d[:, sx - 1] = np.abs(np.sqrt((d[:, sx - 1] - d[:, 0] ** 2).astype(complex) / (ne - 1)))
        else:
            d[:, sx - 1] = 0
        for nc in range(2, nchk + 1):
            d[:, nc - 1] = np.abs(d[:, nc - 1] - d[:, 0])
            er[o, nc - 1] = np.mean(d[:, nc - 1] ** 2)
        fn = compare[o] if o < len(compare) else None
        if fn is not None:                                  # xsim.m:110-143
            esum = np.sum(d[:, 1:sx] ** 2, axis=1) + 1e-20
            ept = (esum > tol) & (np.abs(d[:, 0]) > cutoff)
            dif0 = _compare_values(q, fn, shapes[o])
            ept = ept & (np.abs(dif0) > cutoff)
            mx = np.max(np.abs(dif0)) ** 2
            if not (q.get("relerr", 1) and mx > tol):
                mx = 1.0
            dif = (d[:, 0] - dif0) ** 2
            er[o, sx] = np.mean(dif) / mx
            if ne > 1:
                er[o, sx + 1] = np.sum(ept * (dif / esum))
            er[o, sx + 2] = np.sum(ept)
        er[o, 1:nchk] = er[o, 1:nchk] / mx
        er[o, sx - 1] = np.mean(d[:, sx - 1] ** 2) / mx
        seg = er[o, 1:sx + 1]
        er[o, 0] = np.sum(seg) / (1e-99 + np.sum(seg > 0))
        er[o, :sx + 1] = np.sqrt(er[o, :sx + 1])
        res = np.zeros((d.shape[0], 3))
        res[:, 0] = d[:, 0]
        res[:, 1] = np.sqrt(np.sum(d[:, 1:nchk] ** 2, axis=1)) if nchk > 1 else 0.0
        res[:, 2] = d[:, sx - 1]
        outd.append(res)
    return er, outd


def error_vector(q, er):
    """The error vector of xsim.m:174-188 from the RMS rows of every output (of
    every problem of a sequence)."""
    nchk = 1 + q["checks"][0]
    sx = max(3, 1 + nchk)
    tol = float(q.get("tol", 1e-20))
    error = np.zeros(max(6, sx + 3))
    for c in range(2, sx + 2):
        pts = 1e-99 + np.sum(er[:, c - 1] > tol)
        error[c - 1] = math.sqrt(np.sum(er[:, c - 1] ** 2) / pts)
    error[0] = math.sqrt(np.sum(error[1:sx + 1] ** 2) / (1e-99 + np.sum(error[1:sx + 1] > tol)))
    error[sx + 1:sx + 3] = np.sum(er[:, sx + 1:sx + 3], axis=0)   # chi-squared per point, points
    error[sx + 1] = error[sx + 1] / (1e-20 + error[sx + 2])
    return error


def output_shapes(prob):
    """The shape of every output of a problem ([lines, points, (sites)])."""
    zero = prob.zero_av()
    return [np.asarray(out(zero, prob.q)).shape for out in prob.outputs]


def sequence_inputs(plist):
    """The problems of a sequence as xpreferences.m:11-13,29-32,197 prepares
    them: checks from the first, each default time origin where the previous
    problem ended, the random stream of problem s keyed by s (gpu.seq), the
    ensembles of the first for every problem (xensemble.m:25-30 loops over
    them).  Returns the problems as handed to Problem and their preferences."""
    ps, qs = [], []
    origin = 0.0
    for s, pin in enumerate(plist):
        pk = dict(pin)
        pk["gpu"] = dict(pk.get("gpu") or {}, seq=s)
        if s > 0:
            pk["checks"] = list(qs[0]["checks"])
            if pk.get("ensembles") is not None and \
                    [int(e) for e in _vec(pk["ensembles"], 3, 1)] != qs[0]["ensembles"]:
                raise GpuFormError("the problems of a sequence share the ensembles of the first")
            pk["ensembles"] = list(qs[0]["ensembles"])
        if pk.get("origins") is None:
            q0 = preferences(pk)
            pk["origins"] = [origin] + list(q0["origins"][1:])
        q = preferences(pk)
        origin = q["origins"][0] + q["ranges"][0]
        ps.append(pk)
        qs.append(q)
    return ps, qs


def _to_cells(q, b):
    """A path's final fields [E, sites] (x fastest) as the reference's field
    cell [1, nx, (ny), E] (p.d.ca), for p.transfer."""
    E = q["ensembles"][0]
    dims = space_dims(q)
    b = np.asarray(b).reshape(E, -1)
    if not dims:
        return b.T[None]
    x = _unflatten_sites(b, 1, dims)                     # [E, nx, (ny)]
    return np.moveaxis(x, 0, -1)[None]


def _from_cells(q, a):
    E = q["ensembles"][0]
    dims = space_dims(q)
    shape = (1,) + tuple(dims) + (E,)
    a = np.broadcast_to(np.asarray(a, dtype=complex), shape)[0]
    a = np.moveaxis(a, -1, 0)                              # [E, nx, (ny)]
    if len(dims) == 2:
        a = np.transpose(a, (0, 2, 1))                     # [E, ny, nx]: x fastest
    return np.ascontiguousarray(a).reshape(E, -1)


def xsim_sequence(plist):
    """xsim(p1, p2, ...) (xsim.m:33-38, xensemble.m:25-108): per parallel
    ensemble, sub-ensemble j and check nc the problems run in order on the GPU
    path, each after the first starting from p.transfer(b, v, p) of the
    previous problem's final fields b (default: b itself; v is not available on
    the host and is passed as None); the moments of every problem as
    xensemble accumulates them, and one error vector over the outputs of all
    problems.  Returns (error, data per problem, preferences per problem)."""
    t0 = time.time()
    ps, qs = sequence_inputs(plist)
    if any(q["fields"] != 1 for q in qs):
        raise GpuFormError("sequences of projects with several field components are not on the GPU path")
    probs = []
    try:
        for pk in ps:
            probs.append(make_problem(pk))
        q1 = qs[0]
        e2, e3, nchk = q1["ensembles"][1], q1["ensembles"][2], 1 + q1["checks"][0]
        data = [None] * len(ps)
        for npar in range(1, e3 + 1):
            avs = [dict() for _ in ps]
            for j in range(1, e2 + 1):
                for nc in range(1, nchk + 1):
                    b = None
                    for s, (pk, q, prob) in enumerate(zip(ps, qs, probs)):
                        a0 = None
                        if s > 0:
                            tr = pk.get("transfer")
                            prev = _to_cells(qs[s - 1], b)
                            a0 = _from_cells(q, tr(prev, None, Params(q)) if callable(tr) else prev)
                        av, b = prob.xpath(nc, npar, j, a0=a0, want_final=True)
                        avs[s][(j - 1, nc - 1)] = av
            for s, (q, prob) in enumerate(zip(qs, probs)):
                dp = ensemble_moments(q, None, prob.outputs, None, split=lambda j, nc, _s=s: avs[_s][(j, nc)])
                data[s] = dp if data[s] is None else [x + y for x, y in zip(data[s], dp)]
        rows, outd = [], []
        for q, d, prob in zip(qs, data, probs):
            er, od = statistics_rows(q, d, output_shapes(prob))
            rows.append(er)
            outd.append(od)
        error = error_vector(q1, np.concatenate(rows, axis=0))
        error[-1] = time.time() - t0
    finally:
        for prob in probs:
            prob.close()
    return error, outd, qs


def xsim(p):
    """[error, data, input, raw] = xsim(p) on the GPU path (xsim.m:28-188, with
    the comparisons of p.compare in x space).  data[o] is [ndata, 3] (mean,
    step error, sampling error);
    raw is raw[nc][j] with j over all ensembles(2) x ensembles(3) paths
    (xsim.m:42,56-57), or None without p.rawdata.  A non-blank p.file saves
    error, data, input and raw to that '.mat' file (xsim.m:207-210).  A list of
    problems is a sequence (xsim_sequence): data and input are then per problem
    and raw is None."""
    if isinstance(p, (list, tuple)):
        error, outd, qs = xsim_sequence(p)
        return error, outd, qs, None
    t0 = time.time()
    prob = make_problem(p)
    q = prob.q
    raw = [[] for _ in range(1 + q["checks"][0])] if q["rawdata"] else None
    try:
        data = None
        for npar in range(1, q["ensembles"][2] + 1):               # parfor == serial (xsim.rst:275)
            if raw is not None:
                dp, rp = prob.xensemble(npar, want_raw=True)
                for nc in range(len(raw)):
                    raw[nc].extend(rp[nc])                          # j + e2 (npar - 1)
            else:
                dp = prob.xensemble(npar)
            data = dp if data is None else [a + b for a, b in zip(data, dp)]
        error, outd = statistics(q, data, output_shapes(prob))
        error[-1] = time.time() - t0
        if q["file"].strip():
            from .xfile import xsave
            xsave(q["file"], error, matlab_data(q, prob.zero_av(), prob.outputs, outd), q, raw)
    finally:
        prob.close()
    return error, outd, q, raw


class BatchProblem:
    """Many independent projects in one graph launch (uneven lattices, ensembles
    and drift parameters); each keeps the averages of its own Problem."""

    def __init__(self, problems, autotune=False, serial=False):
        self.qs = [preferences(p) for p in problems]
        self.batch = _lib.Batch([make_desc(q) for q in self.qs], autotune=autotune, serial=serial)
        self.infos = self.batch.infos

    def run(self, npar=1):
        self.batch.run(npar)

    def synchronize(self):
        return self.batch.synchronize()

    def read_averages(self, i):
        return self.batch.read_averages(i)

    def plan(self):
        return self.batch.plan()

    def close(self):
        self.batch.close()
