"""Oracle callables for the named built-in forms of the GPU path (test infrastructure only).

In the reference a user writes ``p.deriv``, ``p.linear``, ``p.initial`` and
``p.observe`` as MATLAB function handles (e.g. Examples/SDE_1D/Kubo.m:16-17,
Examples/SPDE_2D/Soliton.m:16-18, Examples/SPDE_3D/GPE.m:24-25).  The GPU
path offers a library of such handles selected by name through ``p['gpu']``
(precedent: Applications/Quantum/qpreferences.m:48-60 picks built-in derivs by
flag).  This module turns the same ``gpu`` dictionary into FP64 callables with
the reference's calling conventions, so the oracle integrates exactly the
equation the GPU integrates.

Forms (``gpu`` keys):

deriv    'zero'                        D = 0
         'kubo'   c                    D = i c w0 a                       (Kubo.m:17)
         'gp'     g, u, s              D = g |a|^2 a + u a + s (w0 + i w1)   [noises >= 2]
                                       D = g |a|^2 a + u a + s w0            [noises == 1]
         'matrix' M, g, S              D_i = sum_j M_ij a_j + g_i |a_i|^2 a_i + sum_k S_ik w_k
                                       (p.fields coupled SDE components; l0 and c0 per component)
linear   l0, l1[d], l2[d]              L = l0 + sum_d l1[d] D_d + l2[d] D_d^2  (xpropfactor.m:37-39)
initial  'const'  c0                   a = c0
         'sech'   amp, width, eps, center
                                       a = amp sech((x-center)/width) + eps (v0 + i v1)/sqrt2
         'gauss'  amp, width, width_y  a = amp exp(-x^2/(2 width^2) - y^2/(2 width_y^2))
                                       (width_y defaults to width)
         'phase'  amp                  a = amp exp(2 pi i u0),  u0 = uniform(0,1)
         'coherent' alpha              a = alpha + (v0 + i v1)/2  (Wigner vacuum)
observe  'a', 'abs2', 'a_ave', 'abs2_ave', 'abs2_int', 'abs4_ave', 'corr'
         (of component ofield[o] - 1 of a multi-component field, default the first)
"""
import math

import numpy as np

DERIVS = ("zero", "kubo", "gp", "matrix")
INITIALS = ("const", "sech", "gauss", "phase", "coherent")
OBSERVES = ("a", "abs2", "a_ave", "abs2_ave", "abs2_int", "abs4_ave", "corr")


def _space_axes(p):
    return tuple(range(1, p.dimensions))


def make_deriv(g):
    kind = g.get("deriv", "zero")
    if kind == "zero":
        return lambda a, w, p: 0 * a
    if kind == "kubo":
        c = float(g.get("c", 1.0))
        return lambda a, w, p: 1j * c * w[0:1] * a
    if kind == "gp":
        gg = complex(g.get("g", 0.0))
        u = complex(g.get("u", 0.0))
        s = float(g.get("s", 0.0))

        def deriv(a, w, p):
            # noise components: the Gaussian cell, then the k-space filtered cell
            nw = p.noises + p.get("knoises", 0)
            d = gg * (np.conj(a) * a) * a + u * a
            if s != 0.0 and nw >= 2:
                d = d + s * (w[0:1] + 1j * w[1:2])
            elif s != 0.0 and nw == 1:
                d = d + s * w[0:1]
            return d
        return deriv
    if kind == "matrix":
        M = np.asarray(g.get("M"), complex)
        F = M.shape[0]
        gn = np.broadcast_to(np.asarray(g.get("g", 0.0), complex), (F,)).reshape(F, 1)
        S = np.asarray(g.get("S", np.zeros((F, 1))), float).reshape(F, -1)

        def deriv(a, w, p):
            d = np.tensordot(M, a, axes=(1, 0)) + gn * (np.conj(a) * a) * a
            nw = min(S.shape[1], w.shape[0]) if np.ndim(w) else 0
            if nw:
                d = d + np.tensordot(S[:, :nw], w[:nw], axes=(1, 0))
            return d
        return deriv
    raise ValueError("unknown drift form %r" % (kind,))


def make_linear(g):
    lin = g.get("linear", None)
    if lin is None:
        return None
    if g.get("deriv") == "matrix":                      # one l0 per component (an SDE)
        l0v = np.asarray(lin.get("l0", 0.0), complex).reshape(-1, 1)
        return lambda p: l0v
    l0 = complex(lin.get("l0", 0.0))
    l1 = lin.get("l1", 0.0)
    l2 = lin.get("l2", 0.0)

    def linear(p):
        nd = p.dimensions
        c1 = np.broadcast_to(np.asarray(l1, complex), (max(1, nd - 1),))
        c2 = np.broadcast_to(np.asarray(l2, complex), (max(1, nd - 1),))
        L = l0 + 0j
        for d in range(1, nd):
            L = L + c1[d - 1] * p.D[d + 1] + c2[d - 1] * p.D[d + 1] ** 2
        return L
    return linear


def make_initial(g):
    ini = g.get("initial", "const")
    if ini == "const":
        if np.ndim(g.get("c0", 1.0)) > 0:             # one constant per component
            cv = np.asarray(g.get("c0"), complex).reshape(-1, 1)
            return lambda *args: cv + 0j
        c = complex(g.get("c0", 1.0))
        return lambda *args: c + 0 * args[0]
    if ini == "sech":
        amp = float(g.get("amp", 1.0))
        wd = float(g.get("width", 1.0))
        eps = float(g.get("eps", 0.0))
        x0 = float(g.get("center", 0.0))

        def initial(v, p):
            x = p.x if p.dimensions > 1 else 0.0         # an SDE has no space grid: x = 0
            a = amp / np.cosh((x - x0) / wd) + 0j
            if eps != 0.0:
                a = a + eps * (v[0:1] + 1j * v[1:2]) / np.sqrt(2.0)
            return a
        return initial
    if ini == "gauss":
        amp = float(g.get("amp", 1.0))
        wd = float(g.get("width", 1.0))
        wy = float(g.get("width_y", 0.0)) or wd

        def initial(v, p):
            r2 = (p.x ** 2 if p.dimensions > 1 else 0.0) / wd ** 2
            if p.dimensions > 2:
                r2 = r2 + p.y ** 2 / wy ** 2
            return amp * np.exp(-0.5 * r2) + 0j
        return initial
    if ini == "phase":
        amp = float(g.get("amp", 1.0))

        def initial(*args):
            p = args[-1]
            u = args[-2] * p.dv                                  # uniform(0,1): xrandom.m:32 scales by 1/dv
            return amp * np.exp(2j * np.pi * u[0:1])
        return initial
    if ini == "coherent":
        alpha = complex(g.get("alpha", 1.0))

        def initial(v, p):
            return alpha + (v[0:1] + 1j * v[1:2]) / 2.0
        return initial
    raise ValueError("unknown initial form %r" % (ini,))


def lowest_modes(nx, ny, count):
    """FFT indices of the `count` lattice modes of smallest |m|^2, ties broken by
    the angle atan2(my, mx) (the feature modes of the control variate)."""
    cand = sorted((a * a + b * b, math.atan2(b, a), a, b)
                  for a in range(-(nx // 2), nx - nx // 2) for b in range(-(ny // 2), ny - ny // 2))
    return [(a % nx, b % ny) for _, _, a, b in cand[:count]]


def make_observe(name, g=None, field=0):
    if field:
        f0 = make_observe(name, g)
        return lambda a, p: f0(a[field:field + 1], p)
    if name == "lowk":
        # occupation of the feature modes other than k = 0:
        # sum_k (|a_k|^2 - 1/2), a_k = sqrt(dv/Ns) sum_x a(x) exp(-i k.x)
        def f(a, p):
            nx, ny = p.space[0], (p.space[1] if len(p.space) > 1 else 1)
            m = g["modes"]
            modes = lowest_modes(nx, ny, m) if isinstance(m, int) else [tuple(t) for t in m]
            A = np.fft.fftn(a[0:1], axes=_space_axes(p)) * math.sqrt(p.dv / p.nspace)
            A = A.reshape((1, nx, ny) + A.shape[-1:]) if A.ndim == 3 else A
            r = sum(np.abs(A[0, kx, ky]) ** 2 - 0.5 for kx, ky in modes if (kx, ky) != (0, 0))
            return np.asarray(r)[None, :]
        return f
    if name == "abs2k":
        # |a_k|^2 per k site: the xdata k-space path (xnft.m) applies the fft and kfact
        # because apply_forms flags every space dimension in p.transforms
        return lambda a, p: (np.conj(a) * a)[0:1]
    if name == "a":
        return lambda a, p: a[0:1]
    if name == "abs2":
        return lambda a, p: (np.conj(a) * a)[0:1]
    if name == "a_ave":
        return lambda a, p: np.mean(a[0:1], axis=_space_axes(p)) if p.dimensions > 1 else a[0:1]
    if name == "abs2_ave":
        return lambda a, p: (np.mean((np.conj(a) * a)[0:1], axis=_space_axes(p))
                             if p.dimensions > 1 else (np.conj(a) * a)[0:1])
    if name == "abs4_ave":
        return lambda a, p: (np.mean((np.abs(a) ** 4)[0:1], axis=_space_axes(p))
                             if p.dimensions > 1 else (np.abs(a) ** 4)[0:1])
    if name == "abs2_int":
        def f(a, p):
            r = (np.conj(a) * a)[0:1]
            for d in range(p.dimensions - 1, 0, -1):
                r = np.sum(r, axis=d) * p.dx[d]
            return r
        return f
    if name == "corr":
        def f(a, p):
            ax = _space_axes(p)
            A = np.fft.fftn(a[0:1], axes=ax)
            return np.fft.ifftn(np.abs(A) ** 2, axes=ax) / p.nspace
        return f
    raise ValueError("unknown observe form %r" % (name,))


def apply_forms(p_in):
    """Return a copy of the reference-style ``p`` with callables from ``p['gpu']``."""
    p = dict(p_in)
    g = p["gpu"]
    p["deriv"] = make_deriv(g)
    p["linear"] = make_linear(g)
    p["initial"] = make_initial(g)
    names = g.get("observe", ["a"])
    of = list(g.get("ofield") or [1] * len(names))
    p["observe"] = [make_observe(o, g, int(f) - 1) for o, f in zip(names, of)]
    if "abs2k" in names:
        nd = p["dimensions"]
        given = p.get("transforms") or [0] * len(names)
        tflag = [int(np.ravel(t)[0]) if np.size(t) else 0 for t in given] + [0] * len(names)
        p["transforms"] = [[tflag[i]] + [1 if o == "abs2k" else 0] * (nd - 1) for i, o in enumerate(names)]
    return p
