"""FP64 restatement of xSPDE's ensemble integration path (test infrastructure only).

Array convention: a field cell is ``a[field, x, (y,) traj]`` (the reference's
``p.d.ca = [fields, space..., ens]``, xpreferences.m:145).  Noise and random
cells are ``w[comp, x, (y,) traj]`` (xpreferences.m:153-158).  Only one field
cell is supported (``p.fieldcells = 1``), which covers every example on the
path; its ``p.fields`` components may each have their own boundary types and
values (``p.boundaries{2}`` with one row per field, spdpreferences.m:47-66).

Functions and the reference code they restate:

====================  ===========================================  =======================
function              reference                                     pinned by
====================  ===========================================  =======================
xpreferences          Simulator/xpreferences.m:42-219               Batchtest rows (below),
                      (left-out entries: xprefer.m:29-45)           default-padding test
                      Applications/SPDE/spdpreferences.m:16-72
                      Applications/SPDE/xgrid.m:18-48
xpropfactor           Applications/SPDE/xpropfactor.m:23-58       linear-only exactness,
                                                                    Batchtest rows
xprop / xtransform    xprop.m:17-46, xtransform.m:21-70          transforms brute force,
                                                                    Batchtest rows
xsetbound             xsetbound.m:30-59 (boundval constants or     Batchtest #21 (NN),
                      boundfun(a, c, d, p) at p.t; per field)       #42 (Peregrine, as 4
                                                                    fields), #33, heat
                                                                    closed forms
xpatch, da            xpatch.m:24-62, xpropfactor.m:40-51           inhomogeneous heat
                                                                    closed forms (DD, NN,
                                                                    DN, ND), Batchtest #42
                                                                    (time-dependent values)
RK4                   Applications/SDE/RK4.m:7-42                   Batchtest #4, #5, #8,
                                                                    #14, #16-#19, #21, #23
MP                    Applications/SDE/MP.m:10-27                   order test (dt vs dt/2),
                                                                    Kubo closed form,
                                                                    Batchtest #33, #42
xnoise / xrandom      Simulator/xnoise.m:26-60, xrandom.m:18-34     moments, KS test;
  (incl. knoises,     xift.m                                        filtered-noise covariance
  krandoms)                                                         closed form; MATLAB
                                                                    stream itself: parity
                                                                    unpinned
xpath                 Simulator/xpath.m:16-107                      Batchtest rows, checks
                      (incl. time-domain spectra, xpath.m:45-68,    pairing test; spectra:
                      90-106, xpreferences.m:199-203)               geometric-series closed
                                                                    form, OU spectrum
xdata                 Simulator/xdata.m:21-85 (mean branch)       Batchtest rows
xensemble             Simulator/xensemble.m:12-108                  Batchtest #21 (order
                                                                    extrapolation)
xsim statistics       Simulator/xsim.m:64-188                       Batchtest rows,
                                                                    synthetic sigma test
xsim sequences        xsim.m:33-38, xensemble.m:59-63 (transfer),   Batchtest #33
                      xpreferences.m:29-32,197                      (Boundaries2), #20
                                                                    (SolitonDerivS)
xcompress             Applications/SPDE/xcompress.m:18-69           Batchtest #16
Ave / Int / D1        Applications/SPDE/Ave.m, Int.m, D1.m          Batchtest #14, #16-#18,
                                                                    #21, #23
D2                    Applications/SPDE/D2.m:38-78                  Batchtest #33, #20,
                                                                    quadratic with
                                                                    ghost-point ends
xnft                  Applications/SPDE/xnft.m                      Batchtest #16
====================  ===========================================  =======================

Built-in drift/initial/observe *forms* (the named forms the GPU path
implements) are in ``oracle.forms``; they are user-level functions in the
reference's terms (``p.deriv``, ``p.initial``, ``p.observe``).
"""
import math
import time as _time

import numpy as np

from . import philox
from .transforms import transform


class P(dict):
    """Parameter structure with attribute access (the reference's struct ``p``)."""

    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError as e:
            raise AttributeError(k) from e

    def __setattr__(self, k, v):
        self[k] = v

    def copy(self):
        return P(dict.copy(self))


def _listify(x, n=None, default=None):
    if x is None:
        x = default
    if callable(x) or not isinstance(x, (list, tuple)):
        x = [x]
    x = list(x)
    if n is not None:
        while len(x) < n:
            x.append(x[-1] if x else default)
    return x


def _prefer(x, dflt):
    """xprefer.m:29-45: the given entries, then the default's entries for every
    place the input leaves out (dflt is the full default list)."""
    if x is None:
        return list(dflt)
    if callable(x) or not isinstance(x, (list, tuple, np.ndarray)):
        x = [x]
    x = list(x)
    return x + list(dflt[len(x):])


# ----------------------------------------------------------------- preferences
def xpreferences(p_in):
    """Defaults, lattice, scales and grids (xpreferences.m, spdpreferences.m, xgrid.m)."""
    p = P(p_in)
    nd = max(1, int(p.get("dimensions", 1)))
    p.dimensions = nd
    p.fields = int(p.get("fields", 1))
    noises = p.get("noises", None)
    if isinstance(noises, (list, tuple)):                 # the older [x-noises, k-noises] vector
        if len(noises) > 1 and p.get("knoises", None) is None:
            p["knoises"] = noises[1]
        noises = noises[0] if noises else None
    # which cells the project names, before the defaults below fill them in
    given = {k: p.get(k, None) is not None for k in ("knoises", "unoises", "inrandoms", "urandoms", "krandoms")}
    p.knoises = int(p.get("knoises", 0) or 0)
    p.krandoms = int(p.get("krandoms", 0) or 0)
    if noises is None and not given["knoises"] and not given["unoises"]:
        p.noises = p.fields                                             # xpreferences.m:65-70
    else:
        p.noises = int(noises or 0)
    if not given["inrandoms"] and not given["urandoms"] and not given["krandoms"]:
        p.inrandoms = p.noises                                          # xpreferences.m:71-76
    else:
        p.inrandoms = int(p.get("inrandoms", 0) or 0)
    # k-space filters, default identity (xpreferences.m:122-123,419-425)
    p.nfilter = p.get("nfilter", None) or (lambda w, p: w)
    p.rfilter = p.get("rfilter", None) or (lambda w, p: w)
    p.urandoms = int(p.get("urandoms", 0) or 0)
    p.ensembles = list(p.get("ensembles", [1, 1, 1])) + [1] * 3
    p.ensembles = [int(e) for e in p.ensembles[:3]]
    p.ranges = np.asarray(_prefer(p.get("ranges"), [10.0] * nd), float)[:nd]      # xpreferences.m:81
    st = p.get("steps", None)
    st = [1] * nd if st is None else _listify(st)
    p.steps = np.asarray(st + [1] * (nd - len(st)), int)[:nd]
    pts = p.get("points", None)
    pts = np.asarray(_prefer(pts, [51] + [35] * (nd - 1))[:nd], int)             # xpreferences.m:56
    pts[1:] = 1 + (pts[1:] - 1) * p.steps[1:]                    # xpreferences.m:137
    p.points = pts
    og = p.get("origins", None)
    og = [float(p.get("_origin0", 0.0))] + list(-p.ranges[1:] / 2) if og is None else _listify(og)
    p.origins = np.asarray(list(og) + list(-p.ranges[len(og):] / 2), float)[:nd]
    ch = p.get("checks", None)
    ch = [1] + [0] * (nd - 1) if ch is None else _listify(ch)
    p.checks = np.asarray(list(ch) + [0] * (nd - len(ch)), int)[:nd]
    if p.checks[1:].any():
        raise NotImplementedError("spatial checks are outside the integration path")
    p.iterations = int(p.get("iterations", 4))
    p.order = int(p.get("order", 0))
    p.seed = int(p.get("seed", 0))
    p.relerr = int(p.get("relerr", 1))
    p.tol = float(p.get("tol", 1e-20))
    meth = p.get("method", None)
    if meth is None:
        meth = "RK4" if np.prod(p.ensembles) == 1 else "MP"            # xpreferences.m:101-105
    p.method = meth if isinstance(meth, str) else meth.__name__
    p.ipsteps = int(p.get("ipsteps", 2))                               # xpreferences.m:213-218
    dx = p.ranges / np.maximum(1, pts - 1)
    dx = np.where(pts == 1, 1.0, dx)
    p.dx = dx
    p.dv = float(np.prod(dx[1:]))                                      # xpreferences.m:140
    p.noisefactor = 1.0 / p.dv
    p.knoisefactor = float(np.prod(pts[1:])) / p.dv
    p.dk = 2 * np.pi / (pts * dx)
    p.dk[0] = p.dk[0] / p.steps[0]
    p.kfact = dx / math.sqrt(2 * np.pi)
    p.space = tuple(int(n) for n in pts[1:])
    p.nspace = int(np.prod(pts[1:])) if nd > 1 else 1
    p.dt = float(p.ranges[0] / max(1, pts[0] - 1))                     # xpreferences.m:174
    p.nscale = p.steps[0] * (1 + p.checks[0]) / p.dt                   # xpreferences.m:175
    p.nchk = 1 + int(p.checks.sum())
    p.fbsw = 0
    E = p.ensembles[0]
    p.d_ca = (p.fields,) + p.space + (E,)
    # spdpreferences.m:16-72 and xgrid.m:18-48
    bnd = p.get("boundaries", None)
    F = p.fields
    # per dimension, one [lower, upper] row per field (spdpreferences.m:17,47-66:
    # p.boundaries{c,d} is fields x 2; a single row here applies to every field)
    rows = []
    for d in range(1, nd):
        b = [0, 0]
        if bnd is not None and len(bnd) > d - 1 and bnd[d - 1] is not None:
            b = bnd[d - 1]
        b = np.asarray(b, dtype=int)
        if b.ndim == 1:
            b = np.tile(b, (F, 1))
        if b.shape != (F, 2):
            raise ValueError("boundaries of dimension %d: need [lower, upper] or one row per field" % (d + 1))
        rows.append(b)
    # constant boundary values p.boundval (spdpreferences.m:26-43, xboundfun default 0):
    # per space dimension (lower, upper) - field values at Dirichlet ends,
    # derivatives at Neumann (Robin) ends; one pair for every field or one per field
    bv = p.get("boundval", None)
    fbv = [[(0.0, 0.0)] * (nd - 1) for _ in range(F)]
    for d in range(1, nd):
        if bv is None or len(bv) <= d - 1 or bv[d - 1] is None:
            continue
        v = bv[d - 1]
        per_field = F > 1 and len(v) == F and isinstance(v[0], (list, tuple)) and len(v[0]) == 2
        for f in range(F):
            lo, hi = v[f] if per_field else v
            fbv[f][d - 1] = (lo, hi)
    uniform = all((r == r[0]).all() for r in rows) and all(fbv[f] == fbv[0] for f in range(F)) and \
        (F == 1 or p.get("boundfun", None) is None)
    p.setboundaries = int(any((r != 0).any() for r in rows))
    sh = (1,) + p.space + (1,)
    p.xs = []
    ks_f = [[] for _ in range(F)]
    for d in range(1, nd):
        n = int(pts[d])
        p0 = np.arange(n)
        xc = p.origins[d] + p0 * dx[d]
        shape = [1] * (nd + 1)
        shape[d] = n
        p.xs.append(np.broadcast_to(xc.reshape(shape), sh).copy())
        kc = (p0 - n // 2) * p.dk[d]
        dkp = np.pi / ((n - 1) * dx[d]) if n > 1 else 0.0
        for f in range(F):
            t = int(rows[d - 1][f, 0]) + 2 * int(rows[d - 1][f, 1])
            if t == -3:
                kcp = p0 * dkp
            elif t == -1:
                kcp = np.concatenate([[0.0], np.arange(1, n) - 0.5]) * dkp
            elif t == 0:
                kcp = np.fft.ifftshift(kc)
            elif t == 1:
                kcp = np.concatenate([np.arange(1, n) - 0.5, [0.0]]) * dkp
            elif t == 3:
                kcp = p0 * dkp
            else:
                raise ValueError("Invalid boundary type: [%d,%d]" % tuple(rows[d - 1][f]))
            ks_f[f].append(np.broadcast_to(kcp.reshape(shape), sh).copy())
    # field groups sharing one boundary setup: all fields at once when every field
    # has the same boundaries (the single-field path), else one group per field
    # (xtransform.m:25-70 and xsetbound.m:37-55 loop over the field index)
    groups = range(1) if uniform else range(F)
    p.bgroups = []
    for f in groups:
        g = P()
        g.sl = slice(None) if uniform else slice(f, f + 1)
        g.field = None if uniform else f
        g.btypes = [(int(r[f, 0]) + 2 * int(r[f, 1]), [int(r[f, 0]), int(r[f, 1])]) for r in rows]
        g.bvals = [(complex(lo), complex(hi)) for lo, hi in fbv[f]]
        g.da = [0.0] * (nd - 1)
        p.bgroups.append(g)
    p.btypes = p.bgroups[0].btypes
    p.bvals = p.bgroups[0].bvals
    if uniform:
        p.ks = ks_f[0]
    else:
        p.ks = [np.concatenate([ks_f[f][d] for f in range(F)], axis=0) for d in range(nd - 1)]
    names = ["x", "y", "z"]
    for i, (xg, kg) in enumerate(zip(p.xs[:3], p.ks[:3])):
        p[names[i]] = xg
        p["k" + names[i]] = kg
        p["D" + names[i]] = 1j * kg
    p.D = [None, None] + [1j * kg for kg in p.ks]       # p.D{d} = i k_d, d = 2.. (xpropfactor.m:30-33)
    p.r = [None, None] + list(p.xs)                     # p.r{d}, d = 2..: x grids (xgrid.m:19-21)
    p.xc = [p.origins[0] + p.dx[0] * np.arange(pts[0])] + \
           [p.origins[d] + np.arange(pts[d]) * dx[d] for d in range(1, nd)]
    p.kc = [None] + [(np.arange(pts[d]) - pts[d] // 2) * p.dk[d] for d in range(1, nd)]
    p.deriv = p.get("deriv", None) or (lambda a, w, p: 0 * a)
    p.transfer = p.get("transfer", None) or (lambda b, *v: b)         # xpreferences.m:94
    p.linear = p.get("linear", None)
    p.initial = p.get("initial", None) or (lambda v, p: 0)
    obs = p.get("observe", None)
    p.observe = _listify(obs) if obs is not None else [lambda a, p: a]
    nobs = len(p.observe)
    # cells of per-output settings: entries left out take the default
    # (xprefer.m:29-45; transforms zeros, xpreferences.m:240)
    tr = p.get("transforms", None)
    tr = [tr] if tr is not None and not isinstance(tr, (list, tuple)) else tr
    p.transforms = [np.asarray(_prefer(t, [0] * nd)[:nd]) for t in _prefer(tr, [0] * nobs)]
    outs = p.get("output", None)
    p.output = _prefer(outs, [None] * nobs)
    p.output = [o if o is not None else (lambda av, p, _n=n: av[_n])
                for n, o in enumerate(p.output)]
    p.compare = _prefer(p.get("compare", None), [None] * len(p.output))          # xpreferences.m:363
    p.axes = _prefer(p.get("axes", None), [None] * len(p.output))
    p.cutoff = float(p.get("cutoff", -1e20))
    # time-frequency grid of the spectra (xpreferences.m:199-203) and the
    # spectrum switch (xpreferences.m:99,246)
    p.dw = 2 * np.pi / (p.ranges[0] + p.dt)
    p.fsc = math.sqrt(2 * np.pi) / p.dw
    p.plotpts = np.arange(pts[0]) - pts[0] // 2
    p.w = p.plotpts * p.dw
    p.spectrum = int(bool(p.get("spectrum", 0)) or any(t[0] for t in p.transforms))
    p.t = float(p.origins[0])
    p.dtr = 0.0
    return p


# ------------------------------------------------------------------ propagator
def xpropfactor(p):
    """propagator = exp(L * dtr / ipsteps) (xpropfactor.m:26-39); also the
    Neumann patch correction da_d = -(L2 - L1) dt range_d^2 / pi^2 from the first
    two k values of each non-periodic dimension (xpropfactor.m:40-51), per field
    group into g.da (p.da for the first group).  L may carry one row per field
    (p.Dx does when fields have different boundaries)."""
    p.bgroups = [g.copy() for g in p.bgroups]
    for g in p.bgroups:
        g.da = [0.0] * len(g.btypes)
    p.da = p.bgroups[0].da
    if p.linear is None or (not callable(p.linear) and p.linear == 0):
        return None
    L = np.asarray(p.linear(p))
    dt = p.dtr / p.ipsteps
    if p.setboundaries:
        Lf = np.broadcast_to(L, (L.shape[0] if L.ndim == len(p.space) + 2 else 1,) + p.space + (1,))
        for g in p.bgroups:
            Lb = Lf if Lf.shape[0] == 1 else Lf[g.sl]
            for d in range(1, len(g.btypes) + 1):
                i1 = [0] * Lb.ndim
                i2 = list(i1)
                i2[d] = 1
                g.da[d - 1] = -(Lb[tuple(i2)] - Lb[tuple(i1)]) * dt / np.pi ** 2 * p.ranges[d] ** 2
        p.da = p.bgroups[0].da
    return np.exp(L * dt)


def boundary_line(v, a_shape, d):
    """One boundary value as an array over the lattice with axis d of length 1
    (xboundfun.m's [1, .., 1 at d, .., E] shape): a scalar, a vector over the
    other space dimension, or an array of the lattice's space shape (values
    taken at the first site along d)."""
    v = np.asarray(v, dtype=complex)
    nsp = len(a_shape) - 2
    shape = [1] * len(a_shape)
    if v.size == 1:
        return v.reshape(shape)
    space = tuple(a_shape[1:1 + nsp])
    if v.size == int(np.prod(space)):
        v = np.take(v.reshape(space), [0], axis=d - 1)
    other = [k for k in range(1, 1 + nsp) if k != d]
    if len(other) != 1 or v.size != a_shape[other[0]]:
        raise ValueError("boundary value of size %d does not fit dimension %d" % (v.size, d))
    shape[other[0]] = v.size
    return v.reshape(shape)


def boundary_values(a, p, d, g=None, shape=None):
    """(lower, upper) boundary values of space dimension d at time p.t for field
    group g (default: the first): p.boundfun(a, c, d, p) (reference numbering:
    c = 1, d = 2, 3 for x, y; xpatch.m:37, xsetbound.m:39) when given - one
    (lower, upper) pair, or one per field like the reference's fields x 2 cell -
    else the constants p.boundval (spdpreferences.m:26-43; default 0).  The
    values are shaped for the group's part of the field (``shape``)."""
    g = g or p.bgroups[0]
    shape = shape or a.shape
    f = p.get("boundfun", None)
    if f is None:
        lo, hi = g.bvals[d - 1]
    else:
        b = f(a, 1, d + 1, p)
        if p.fields > 1 and len(b) == p.fields and isinstance(b[0], (list, tuple)):
            b = b[g.field if g.field is not None else 0]
        lo, hi = b[0], b[1]
    return boundary_line(lo, shape, d), boundary_line(hi, shape, d)


def _by_group(p, a, fn):
    """Apply fn(sub_array, group) to every field group and stack the results
    along the field axis (one group: the whole field at once)."""
    if len(p.bgroups) == 1:
        return fn(a, p.bgroups[0])
    parts = [fn(a[g.sl], g) for g in p.bgroups]
    if isinstance(parts[0], tuple):
        return tuple(np.concatenate([np.broadcast_to(q[k], a[g.sl].shape) for q, g in zip(parts, p.bgroups)],
                                    axis=0) for k in range(len(parts[0])))
    return np.concatenate(parts, axis=0)


def xpatch(a, p):
    """Boundary patch (xpatch.m:24-62): returns (a - pa, pa + lp) where pa is the
    per-dimension polynomial matching the boundary values and lp the Neumann
    (type -3) correction db * da_d, per field group."""
    a_full = a

    def one(a, g):
        pa = np.zeros(a.shape, dtype=complex)
        lp = 0.0
        for d, (t, _) in enumerate(g.btypes, start=1):
            if t == 0 or a.shape[d] <= 1:
                continue
            b1, b2 = boundary_values(a_full, p, d, g, a.shape)
            rng = p.ranges[d]
            x = p.xs[d - 1] - p.origins[d]
            if t == -3:
                db = (b2 - b1) / rng
                pa = pa + b1 * x + 0.5 * x ** 2 * db
                lp = lp + db * g.da[d - 1]
            elif t == -1:
                pa = pa + b1 + x * b2
            elif t == 1:
                pa = pa + b2 + (x - rng) * b1
            elif t == 3:
                pa = pa + b1 + x * (b2 - b1) / rng
        return a - pa, pa + lp
    return _by_group(p, a, one)


def xsetbound(a, p):
    """Dirichlet ends set to their boundary values, dimension by dimension (later
    dimensions win at corners), per field group (xsetbound.m:30-59): p.boundfun
    at p.t, else the p.boundval constants (default 0, spdpreferences.m:26-43,75-93)."""
    if not p.setboundaries:
        return a
    a_full = np.array(a, dtype=complex)

    def one(a, g):
        a = np.array(a, dtype=complex)
        for d, (t, b) in enumerate(g.btypes, start=1):
            if t == 0:
                continue
            lo, hi = boundary_values(a_full, p, d, g, a.shape)
            if b[0] == 1:
                idx = [slice(None)] * a.ndim
                idx[d] = slice(0, 1)
                a[tuple(idx)] = np.broadcast_to(lo, a[tuple(idx)].shape)
            if b[1] == 1:
                idx = [slice(None)] * a.ndim
                idx[d] = slice(a.shape[d] - 1, a.shape[d])
                a[tuple(idx)] = np.broadcast_to(hi, a[tuple(idx)].shape)
        return a
    return _by_group(p, a_full, one)


def xtransform(a, p, inverse):
    """Per space dimension spectral transform of xtransform.m:33-70, with each
    field's own boundary type."""
    def one(a, g):
        for d, (t, _) in enumerate(g.btypes, start=1):
            if a.shape[d] > 1:
                a = transform(a, d, t, inverse)
        return a
    return _by_group(p, np.asarray(a), one)


def xprop(a, p, prop):
    """One interaction-picture propagation (xprop.m:17-47): subtract the boundary
    patch, transform, propagate, inverse transform, add the patch back."""
    if prop is not None:
        a = np.asarray(a, complex)
        pat = 0.0
        if p.setboundaries:
            a, pat = xpatch(a, p)
        a = xtransform(a, p, 0)
        a = prop * a
        a = xtransform(a, p, 1) + pat
    if p.setboundaries:
        a = xsetbound(a, p)
    return a


# --------------------------------------------------------------------- methods
def RK4(a, w, p, prop):
    """xSPDE fourth-order interaction-picture Runge-Kutta (RK4.m:7-42)."""
    dt = 0.5 * p.dtr
    t = p.t
    if p.setboundaries:
        a = xsetbound(a, p)
    a1 = a + p.deriv(a, w, p) * dt
    a1 = xprop(a1, p, prop)
    a = xprop(a, p, prop)
    p.t = t + dt
    a2 = a + p.deriv(a1, w, p) * dt
    a2 = xsetbound(a2, p)
    a3 = a + p.deriv(a2, w, p) * 2 * dt
    a3 = xsetbound(a3, p)
    a = (a1 + 2 * a2 + a3 - a) / 3
    p.t = t + 2 * dt
    a = xprop(a, p, prop)
    a3 = xprop(a3, p, prop)
    a = a + p.deriv(a3, w, p) * dt / 3
    a = xsetbound(a, p)
    p.t = t
    return a


def MP(a, w, p, prop):
    """Interaction-picture midpoint with p.iterations fixed-point passes (MP.m:10-27)."""
    dt = p.dtr / 2.0
    t = p.t
    a = xprop(a, p, prop)
    p.t = t + dt
    a0 = a
    for _ in range(p.iterations):
        a = a0 + p.deriv(a, w, p) * dt
        a = xsetbound(a, p)
    a = 2 * a - a0
    p.t = t + 2 * dt
    a = xprop(a, p, prop)
    p.t = t
    return a


METHODS = {"RK4": RK4, "MP": MP}


# ----------------------------------------------------------------------- noise
def _site_major_to_cells(x, comps, space, E):
    """[E, site*comps+comp] -> [comp, x, (y,), E] with site = ix + nx*iy."""
    x = x.reshape((E,) + tuple(reversed(space)) + (comps,))
    perm = (len(space) + 1,) + tuple(range(len(space), 0, -1)) + (0,)
    return np.ascontiguousarray(np.transpose(x, perm))


def xift(a, p):
    """Inverse FFT over the space dimensions (xift.m:8-11)."""
    for d in range(1, p.dimensions):
        a = np.fft.ifft(a, axis=d)
    return a


def noise_draw(p, nsp, q, seq=0, inst=0):
    """One xnoise draw: the Gaussian cell N(0,1)*sqrt(nscale*noisefactor)
    (xnoise.m:29-37), then the k-space cell xift(nfilter(z))*sqrt(nscale*knoisefactor)
    with z ~ N(0,1) per k site in FFT order (xnoise.m:39-50); components stacked."""
    E = p.ensembles[0]
    ns = max(1, p.nspace)
    n = philox.normals(nsp, philox.TAG_NOISE, q, np.arange(E), ns * p.noises, seq, inst)
    w = _site_major_to_cells(n, p.noises, p.space, E) * math.sqrt(p.nscale * p.noisefactor)
    if p.knoises:
        z = philox.normals(nsp, philox.TAG_KNOISE, q, np.arange(E), ns * p.knoises, seq, inst)
        zk = p.nfilter(_site_major_to_cells(z, p.knoises, p.space, E), p)
        wk = xift(zk, p) * math.sqrt(p.nscale * p.knoisefactor)
        w = np.concatenate([w.astype(complex), wk], axis=0) if p.noises else wk
    return w


def xrandom(p, nsp, seq=0, inst=0):
    """Initial randoms (xrandom.m:18-34): Gaussians of variance 1/dv, k-space
    filtered randoms sqrt(knoisefactor) xift(rfilter(z)), then uniforms."""
    E = p.ensembles[0]
    ns = max(1, p.nspace)
    v = []
    if p.inrandoms:
        n = philox.normals(nsp, philox.TAG_GAUSS_INIT, 0, np.arange(E), ns * p.inrandoms, seq, inst)
        v.append(_site_major_to_cells(n, p.inrandoms, p.space, E) * math.sqrt(p.noisefactor))
    if p.krandoms:                                                      # xrandom.m:24-29
        z = philox.normals(nsp, philox.TAG_KRAND_INIT, 0, np.arange(E), ns * p.krandoms, seq, inst)
        zk = p.rfilter(_site_major_to_cells(z, p.krandoms, p.space, E), p)
        v.append(math.sqrt(p.knoisefactor) * xift(zk, p))
    if p.urandoms:
        u = philox.uniforms(nsp, philox.TAG_UNIFORM_INIT, 0, np.arange(E), ns * p.urandoms, seq, inst)
        v.append(_site_major_to_cells(u, p.urandoms, p.space, E) * p.noisefactor)
    if not v:
        v = [np.zeros((1,) + p.space + (E,))]
    return v


# ------------------------------------------------------------------------ data
def xnft(a, trans, p):
    """k-space observable transform (xnft.m): fft along flagged dims times kfact."""
    for d in range(1, p.dimensions):
        if trans[d] > 0 and a.shape[d] > 1:
            a = np.fft.fft(a, axis=d) * p.kfact[d]
    return a


def xdata(a, n, p):
    """Ensemble mean of observable n (xdata.m:25-84, mean branch): [lines, nsp_o]."""
    if np.any(p.transforms[n][1:]):
        a = xnft(a, p.transforms[n], p)
    o = np.asarray(p.observe[n](a, p))
    E = a.shape[-1]
    if o.ndim == 0 or o.shape[-1] != E:
        o = np.broadcast_to(o, tuple(np.shape(o)) + (E,)) if o.ndim == 0 else \
            np.broadcast_to(o[..., None], o.shape + (E,))
    lines = o.shape[0] if o.ndim > 1 else 1
    o = o.reshape(lines, -1, E)
    return o.mean(axis=-1)


# ------------------------------------------------------------------------ path
def xpath(a, nc, p, prop, nsp, seq=0, inst=0, noise_fn=None):
    """Time loop over points x steps x checks (xpath.m:29-86).

    ``noise_fn(q)`` overrides the counter stream: it must return draw q.
    Returns the final field and av[o] of shape [lines, points, nsp_o].
    """
    npts = int(p.points[0])
    nsteps = int(p.steps[0])
    chk = int(p.checks[0])
    method = METHODS[p.method]
    av = [None] * len(p.observe)
    q = 0
    draw = noise_fn or (lambda qq: noise_draw(p, nsp, qq, seq, inst))
    # the field at every step for the time-domain spectra (xpath.m:45-53,64-68):
    # aft[:, n1] after step n1 (n1 = 0: the initial field); a fine run with
    # checks stores the mean of its two half steps' results
    aft = np.zeros((npts * nsteps,) + a.shape, dtype=complex) if p.spectrum else None
    n1 = -1
    for n in range(npts):
        for step in range(nsteps):
            n1 += 1
            if n > 0 or step > 0 or npts == 1:
                w = None
                for c in range(chk + 1):
                    wn = draw(q)
                    q += 1
                    if w is not None and chk and nc == 2:
                        wn = (wn + w) / 2                                 # xnoise.m:33-34
                    w = wn
                    if nc == 1 or c == 1:
                        a = method(a, w, p, prop)
                        if aft is not None:
                            aft[n1] = (a + aft[n1]) / 2 if (c == 1 and nc == 1) else a
                        w = None
                        p.t = p.t + p.dtr
            elif aft is not None:
                aft[n1] = a
            if step == 0:
                for o in range(len(p.observe)):
                    if p.transforms[o][0]:
                        continue                                          # xpath.m:80
                    d = xdata(a, o, p)
                    if av[o] is None:
                        av[o] = np.zeros((d.shape[0], npts, d.shape[1]), dtype=complex)
                    av[o][:, n, :] = d
    if aft is not None:
        # xpath.m:90-105: ifft over time times fsc, fftshift, keep the points(1)
        # frequencies around zero (specpts), observables of the transformed field
        M = npts * nsteps
        ft = np.fft.ifft(aft, axis=0) * p.fsc
        ft = np.fft.fftshift(ft, axes=0)
        specpts = p.plotpts + M // 2
        ft = ft[specpts]
        for o in range(len(p.observe)):
            if not p.transforms[o][0]:
                continue
            for n in range(npts):
                d = xdata(ft[n], o, p)
                if av[o] is None:
                    av[o] = np.zeros((d.shape[0], npts, d.shape[1]), dtype=complex)
                av[o][:, n, :] = d
    return a, av


def initial_field(p, v):
    a = np.asarray(p.initial(*v, p), dtype=complex)
    return np.broadcast_to(a, p.d_ca).copy() if a.shape != p.d_ca else a.copy()


# -------------------------------------------------------------------- ensemble
def run_path(p, j, nc, npar=1, seq=0, inst=0, noise_fn=None, a0=None, prev=None):
    """One (sub-ensemble j, check nc) call of p.path as xensemble.m:26-79 sets it up.
    ``prev``: the final field of the previous element of a sequence, handed to
    p.transfer(b, v, p) for the initial field (xensemble.m:59-63)."""
    p = p.copy()
    nsp = p.seed + npar + (j - 1) * p.ensembles[2]                      # xensemble.m:26
    if p.checks[0] and nc != 2:
        p.dtr = p.dt / (2 * p.steps[0])                                  # xensemble.m:41-45
    else:
        p.dtr = p.dt / p.steps[0]
    prop = xpropfactor(p)
    v = xrandom(p, nsp, seq, inst)
    p.t = float(p.origins[0])
    if a0 is not None:
        a = np.array(a0, dtype=complex)
    elif prev is not None:
        a = np.asarray(p.transfer(prev, *v, p), dtype=complex)
        a = np.broadcast_to(a, p.d_ca).copy() if a.shape != p.d_ca else a.copy()
    else:
        a = initial_field(p, v)
    return xpath(a, nc, p, prop, nsp, seq, inst, noise_fn)


def xensemble(npar, p, path_fn=None):
    """Sub-ensemble and check loops with D, D^2 accumulation (xensemble.m:12-108).

    ``path_fn(p, j, nc, npar)`` may replace the FP64 path (used to feed GPU
    averages through the same bookkeeping).  Returns data[o] of shape
    [ndata_o, m1].
    """
    e2, e3 = p.ensembles[1], p.ensembles[2]
    ne = 1.0 / (e2 * e3) - 1
    m = p.fbsw + p.nchk
    m1 = max(3, m + 1)
    data = [None] * len(p.output)
    for j in range(1, e2 + 1):
        D = {}
        for nc in range(1, p.nchk + 1):
            if path_fn is None:
                _, av = run_path(p, j, nc, npar)
            else:
                av = path_fn(p, j, nc, npar)
            for o, out in enumerate(p.output):
                D[(nc, o)] = np.real(np.asarray(out(av, p))).reshape(-1)
        for o in range(len(p.output)):
            nes = 1 + ne
            if p.checks[0] and p.order > 0:                              # xensemble.m:93-97
                dl = (D[(1, o)] - D[(2, o)]) / (2.0 ** p.order - 1.0)
                D[(2, o)] = D[(1, o)]
                D[(1, o)] = D[(1, o)] + dl
            if data[o] is None:
                data[o] = np.zeros((D[(1, o)].size, m1))
            for n1 in range(1, m + 1):
                data[o][:, n1 - 1] += D[(n1, o)] * nes
            data[o][:, m1 - 1] += D[(1, o)] ** 2 * nes
    return data


def xensemble_seq(npar, ps):
    """xensemble.m:25-108 over a sequence of problems: per (j, nc) the elements run
    in order, each after the first starting from p.transfer of the previous
    element's final field; D, D^2 per element as in xensemble.  Returns one
    data list per element."""
    p1 = ps[0]
    e2, e3 = p1.ensembles[1], p1.ensembles[2]
    nes = 1.0 / (e2 * e3)
    m = p1.fbsw + p1.nchk
    m1 = max(3, m + 1)
    data = [[None] * len(p.output) for p in ps]
    for j in range(1, e2 + 1):
        D = [{} for _ in ps]
        for nc in range(1, p1.nchk + 1):
            b = None
            for s, p in enumerate(ps):
                b, av = run_path(p, j, nc, npar, seq=s, prev=b if s > 0 else None)
                for o, out in enumerate(p.output):
                    D[s][(nc, o)] = np.real(np.asarray(out(av, p))).reshape(-1)
        for s, p in enumerate(ps):
            for o in range(len(p.output)):
                Ds = D[s]
                if p.checks[0] and p.order > 0:                          # xensemble.m:93-97
                    dl = (Ds[(1, o)] - Ds[(2, o)]) / (2.0 ** p.order - 1.0)
                    Ds[(2, o)] = Ds[(1, o)]
                    Ds[(1, o)] = Ds[(1, o)] + dl
                if data[s][o] is None:
                    data[s][o] = np.zeros((Ds[(1, o)].size, m1))
                for n1 in range(1, m + 1):
                    data[s][o][:, n1 - 1] += Ds[(n1, o)] * nes
                data[s][o][:, m1 - 1] += Ds[(1, o)] ** 2 * nes
    return data


# ------------------------------------------------------------------ statistics
def _grid_for_output(p, n, shape, kspace):
    """Broadcast grids p.t, p.x, ... for compare functions (xgrid.m:49-56)."""
    q = p.copy()
    nd = p.dimensions
    full = len(shape) - 2 == nd - 1
    axes = [p.xc[0]]
    for d in range(1, nd):
        if full and shape[d + 1] == 1 < p.points[d]:
            axes.append(p.xc[d][:1])                            # an axis the output integrated out
        elif kspace[d]:
            axes.append(np.fft.fftshift(np.fft.fftfreq(p.points[d], 1.0 / p.points[d])) * p.dk[d])
        else:
            axes.append(p.xc[d])
    grids = []
    for i, ax in enumerate(axes):
        sh = [1] * (nd + 1)
        sh[i + 1] = len(ax)
        grids.append(ax.reshape(sh))
    q.t = grids[0]
    names = ["x", "y", "z"]
    for i in range(1, min(nd, 4)):
        q[names[i - 1]] = grids[i]
        q["k" + names[i - 1]] = grids[i]
    q.r = [None] + grids                                # p.r{1} = t, p.r{d} (xgrid.m:52-55)
    return q, full


def xcompress_idx(axes_o, shape):
    """Index lists per axis for xcompress.m:36-68 (0 all, -1 midpoint, -2 last)."""
    idx = []
    for d, npt in enumerate(shape[1:]):
        ind = 0 if axes_o is None or d >= len(axes_o) else axes_o[d]
        if np.isscalar(ind):
            if ind == 0:
                sel = np.arange(npt)
            elif ind == -1:
                sel = np.array([npt // 2])
            elif ind == -2:
                sel = np.array([npt - 1])
            else:
                sel = np.array([int(ind) - 1])
        else:
            sel = np.asarray(ind, int) - 1
        idx.append(sel)
    return idx


def xsim_stats(p, data, elapsed=0.0, rows=False):
    """Final sampling/step/difference errors of xsim.m:64-188 (no bins/scatters).

    ``data[o]`` is [ndata_o, m1] as returned by xensemble summed over npar.
    Returns (error vector, reduced per-output data list).
    """
    ne = p.ensembles[1] * p.ensembles[2]
    n1 = ne - 1
    nc1 = p.nchk + p.fbsw
    sx = max(3, 1 + nc1)
    nout = len(p.output)
    er = np.zeros((nout, max(6, nc1 + 4)))
    em = np.zeros_like(er)
    outd = []
    for o in range(nout):
        d = np.array(data[o], dtype=float)
        shape = p._avshape[o]                                           # (lines, nt, *space_o)
        kspace = p.transforms[o] if len(shape) - 2 == p.dimensions - 1 else np.zeros(p.dimensions)
        dd = d.reshape(shape + (d.shape[1],))
        for ax in range(1, p.dimensions):
            if kspace[ax] and len(shape) - 2 == p.dimensions - 1:
                dd = np.fft.fftshift(dd, axes=1 + ax)                  # xsim.m:77-81
        idx = xcompress_idx(p.axes[o], shape)
        dd = dd[np.ix_(*([np.arange(shape[0])] + idx + [np.arange(d.shape[1])]))]
        cshape = dd.shape[:-1]
        d = dd.reshape(-1, d.shape[1])
        mx = np.max(np.abs(d[:, 0])) ** 2
        if not (p.relerr and mx > p.tol):
            mx = 1.0
        if ne > 1:
            d[:, sx - 1] = np.abs(np.sqrt((d[:, sx - 1] - d[:, 0] ** 2).astype(complex) / n1))
        else:
            d[:, sx - 1] = 0
        for nc in range(2, nc1 + 1):
            d[:, nc - 1] = np.abs(d[:, nc - 1] - d[:, 0])
            er[o, nc - 1] = np.mean(d[:, nc - 1] ** 2)
            em[o, nc - 1] = math.sqrt(np.max(d[:, nc - 1] ** 2))
        esum = np.sum(d[:, 1:sx] ** 2, axis=1) + 1e-20
        ept = (esum > p.tol) & (np.abs(d[:, 0]) > p.cutoff)
        cmp = p.compare[o] if o < len(p.compare) else None
        if cmp is not None:
            q, full = _grid_for_output(p, o, shape, kspace)
            c = np.asarray(cmp(q), dtype=float)
            if full:
                c = np.broadcast_to(c, shape)
            else:
                c = np.broadcast_to(c, (shape[0], shape[1]) + (1,) * (p.dimensions - 1))
                c = c.reshape(shape[0], shape[1], *([1] * (len(shape) - 2)))
                c = np.broadcast_to(c, shape)
            c = c[np.ix_(*([np.arange(shape[0])] + idx))]
            dif0 = c.reshape(-1)
            ept = ept & (np.abs(dif0) > p.cutoff)
            mx = np.max(np.abs(dif0)) ** 2
            if not (p.relerr and mx > p.tol):
                mx = 1.0
            dif = (d[:, 0] - dif0) ** 2
            em[o, sx] = math.sqrt(np.max(dif) / mx)
            er[o, sx] = np.mean(dif) / mx
            if ne > 1:
                er[o, sx + 1] = np.sum(ept * (dif / esum))
            er[o, sx + 2] = np.sum(ept)
        er[o, 1:p.nchk] = er[o, 1:p.nchk] / mx
        er[o, sx - 1] = np.mean(d[:, sx - 1] ** 2) / mx
        em[o, sx - 1] = math.sqrt(np.max(d[:, sx - 1] ** 2) / mx)
        seg = er[o, 1:sx + 1]
        er[o, 0] = np.sum(seg) / (1e-99 + np.sum(seg > 0))
        er[o, :sx + 1] = np.sqrt(er[o, :sx + 1])
        outd.append(d.reshape(cshape + (d.shape[1],)))
    if rows:
        return er, em, outd
    return _combine_errors(p, er, elapsed), outd


def _stats_rows(p, data):
    """Per-output rows (er, em) and reduced data of xsim.m:64-173 for one element."""
    return xsim_stats(p, data, rows=True)


def _combine_errors(p, er, elapsed=0.0):
    """The error vector of xsim.m:174-188 from the per-output rows of every
    element of the sequence."""
    nc1 = p.nchk + p.fbsw
    sx = max(3, 1 + nc1)
    error = np.zeros(max(6, sx + 3))
    for c in range(2, sx + 2):
        pts = 1e-99 + np.sum(er[:, c - 1] > p.tol)
        error[c - 1] = math.sqrt(np.sum(er[:, c - 1] ** 2) / pts)
    error[0] = math.sqrt(np.sum(error[1:sx + 1] ** 2) / (1e-99 + np.sum(error[1:sx + 1] > p.tol)))
    error[sx + 1:sx + 3] = np.sum(er[:, sx + 1:sx + 3], axis=0)
    error[sx + 1] = error[sx + 1] / (1e-20 + error[sx + 2])
    error[sx + 2] = elapsed
    return error


def xsim(p_in, path_fn=None):
    """[error, data, p] = xsim(p) (xsim.m:28-188).  A list of problems is a
    sequence (xsim.m:33-38, xpreferences.m:11-13,197): checks from the first
    element, each default time origin where the previous element ended, the
    error vector over the outputs of every element; data and p are then lists."""
    if isinstance(p_in, (list, tuple)):
        return _xsim_sequence(p_in)
    t0 = _time.time()
    p = xpreferences(p_in)
    data = None
    for npar in range(1, p.ensembles[2] + 1):                         # parfor == serial, xsim.rst:275
        dp = xensemble(npar, p, path_fn)
        data = dp if data is None else [x + y for x, y in zip(data, dp)]
    # shapes of the averages [lines, points, space_o] per output, from one probe
    if "_avshape" not in p:
        p._avshape = _probe_shapes(p)
    err, outd = xsim_stats(p, data, _time.time() - t0)
    return err, outd, p


def _xsim_sequence(plist):
    t0 = _time.time()
    ps = []
    origin = 0.0
    for k, pin in enumerate(plist):
        q = dict(pin)
        if k > 0:
            q["checks"] = list(ps[0].checks)                         # xpreferences.m:29-32
        q["_origin0"] = origin
        p = xpreferences(q)
        origin = p.origins[0] + p.ranges[0]                           # xpreferences.m:197
        ps.append(p)
    data = None
    for npar in range(1, ps[0].ensembles[2] + 1):
        dp = xensemble_seq(npar, ps)
        data = dp if data is None else [[x + y for x, y in zip(u, w)] for u, w in zip(data, dp)]
    ers, outs = [], []
    for p, d in zip(ps, data):
        p._avshape = _probe_shapes(p)
        er, _, outd = _stats_rows(p, d)
        ers.append(er)
        outs.append(outd)
    err = _combine_errors(ps[0], np.concatenate(ers, axis=0), _time.time() - t0)
    return err, outs, ps


def _probe_shapes(p):
    """Output shapes from observing a zero field (xpreferences.m:262-343)."""
    q = p.copy()
    q.ensembles = [1, 1, 1]
    a = np.zeros((p.fields,) + p.space + (1,), complex)
    av = []
    obs = []
    for o in range(len(p.observe)):
        d = xdata(a, o, q)
        av.append(np.zeros((d.shape[0], int(p.points[0]), d.shape[1]), complex))
        raw = np.asarray(p.observe[o](a, q))                     # [lines, space_o..., E]
        obs.append(tuple(raw.shape[1:-1]) if raw.ndim == p.dimensions + 1 else ())
    shapes = []
    for n, out in enumerate(p.output):
        r = np.asarray(out(av, p))
        lines, nt = r.shape[0], r.shape[1]
        nsp = int(np.prod(r.shape[2:]))
        if nsp == p.nspace and p.dimensions > 1:
            shapes.append((lines, nt) + p.space)
        elif n < len(obs) and 1 < nsp == int(np.prod(obs[n] or (0,))):
            shapes.append((lines, nt) + obs[n])                  # some axes integrated out (size 1)
        else:
            shapes.append((lines, nt))
    return shapes


# ------------------------------------------------------- observable helpers
def Ave(o, p):
    """Lattice average over space axes (Ave.m, default weights)."""
    return np.mean(o, axis=tuple(range(1, p.dimensions)), keepdims=False)


def Int(o, *args):
    """Lattice integral over the space axes (Applications/SPDE/Int.m:23-61, no
    limits): Int(o, p) uses the steps p.dx, Int(o, dx, p) the given steps (p.dk
    for transformed fields); a space dimension with step 0 is not integrated."""
    p = args[-1]
    dx = p.dx if len(args) == 1 else np.asarray(args[0], float)
    partial = bool(np.any(np.asarray(dx[1:p.dimensions]) <= 0))  # MATLAB's sum keeps the other axes
    r = o
    for d in range(p.dimensions - 1, 0, -1):       # last axis first: lower indices stay valid
        if dx[d] > 0:
            r = np.sum(r, axis=d, keepdims=partial) * dx[d]
    return r


def D1(a, d, p):
    """Central first derivative along space axis d (D1.m), Neumann ends = boundval 0,
    per field group."""
    ax = d - 1

    def one(a, g):
        r = (np.roll(a, -1, axis=ax) - np.roll(a, 1, axis=ax)) / (2 * p.dx[d - 1])
        t, b = g.btypes[d - 2]
        if b[0] == -1:
            idx = [slice(None)] * a.ndim
            idx[ax] = 0
            r[tuple(idx)] = 0
        if b[1] == -1:
            idx = [slice(None)] * a.ndim
            idx[ax] = -1
            r[tuple(idx)] = 0
        return r
    return _by_group(p, np.asarray(a), one)


def D2(a, d, p):
    """Central second derivative along space axis d (Applications/SPDE/D2.m:38-78,
    a derivative function: p.indext = 0, cell 1).  Periodic wrap; a field whose
    lower end is not periodic gets end values from its boundary types: at a
    Neumann end the ghost-point form 2((a_1 - a_0)/dx - b_lower)/dx (upper:
    2(b_upper + (a_{N-2} - a_{N-1})/dx)/dx) with the derivative values of
    p.boundval; at any other end the neighbouring interior value (a continuous
    D2).  With a p.boundfun, the values are those stored before the path, at
    t = origins(1) - 1 (xensemble.m:66-72)."""
    ax = d - 1
    dx = p.dx[d - 1]
    a = np.asarray(a)

    def one(a, g):
        r = (np.roll(a, -1, axis=ax) + np.roll(a, 1, axis=ax) - 2 * a) / dx ** 2
        t, b = g.btypes[d - 2]
        if b[0] == 0:
            return r
        q = p.copy()
        q.t = float(p.origins[0]) - 1
        lo, hi = boundary_values(a, q, d - 1, g, a.shape)

        def at(k):
            idx = [slice(None)] * a.ndim
            idx[ax] = slice(k, k + 1) if k != -1 else slice(a.shape[ax] - 1, a.shape[ax])
            return tuple(idx)
        n = a.shape[ax]
        if b[0] == -1:
            r[at(0)] = 2 * ((a[at(1)] - a[at(0)]) / dx - lo) / dx
        else:
            r[at(0)] = r[at(1)]
        if b[1] == -1:
            r[at(-1)] = 2 * (hi + (a[at(n - 2)] - a[at(-1)]) / dx) / dx
        else:
            r[at(-1)] = r[at(n - 2)]
        return r
    return _by_group(p, a, one)
