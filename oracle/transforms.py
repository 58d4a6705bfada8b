"""Spectral transforms of the reference, restated in float64 (test infrastructure only).

Each function acts along one axis of a complex array and follows the named
reference file line by line in meaning (not in code):

=============  =====================================  ==========================
function       reference                               pinned by
=============  =====================================  ==========================
xdct1          Applications/SPDE/xdct1.m:8-20          brute-force sum (xdct1.m code,
                                                       symmetric extension) +
                                                       scipy.fft.dct(type=1) +
                                                       self-inverse
xdct2          xdct2.m:11-29                           brute-force sum (header
                                                       def. on N-1 points) +
                                                       scipy.fft.dct(type=2)
xdct3          xdct3.m:13-33                           brute force + inverse of xdct2
xdst1          xdst1.m:8-19                            brute force + scipy dst(type=1)
                                                       + self-inverse
xdst2          xdst2.m:13-33                           brute force + scipy dst(type=2)
xdst3          xdst3.m:12-35                           brute force + inverse of xdst2
fft / ifft     xtransform.m:45,60 (MATLAB fft)         numpy.fft (textbook DFT)
=============  =====================================  ==========================

Note (xdct1.m:4-5): the header comment writes the end term as
``1/2(a_1 - (-1)^n a_N)``; the code (``xdct1.m:17-20``, symmetric extension
then FFT) gives ``+(-1)^k a_N``.  The code is what runs, so it is followed.
"""
import numpy as np


def _mv(a, axis):
    return np.moveaxis(a, axis, -1)


def _back(a, axis):
    return np.moveaxis(a, -1, axis)


def xdct1(a, axis):
    """DCT-I via FFT of the even extension of length 2(N-1), scaled 1/sqrt(2(N-1))."""
    a = _mv(np.asarray(a), axis)
    n = a.shape[-1]
    if np.iscomplexobj(a):
        return _back(_mv(xdct1(_back(a.real, axis), axis), axis)
                     + 1j * _mv(xdct1(_back(a.imag, axis), axis), axis), axis)
    ext = np.concatenate([a, a[..., n - 2:0:-1]], axis=-1)
    y = np.fft.fft(ext, axis=-1)[..., :n].real * np.sqrt(1.0 / (2 * (n - 1)))
    return _back(y, axis)


def xdct2(a, axis):
    """DCT-II on the first N-1 points (last output fixed to zero), norm sqrt(2/(N-1))."""
    a = _mv(np.asarray(a), axis)
    if np.iscomplexobj(a):
        return _back(_mv(xdct2(_back(a.real, axis), axis), axis)
                     + 1j * _mv(xdct2(_back(a.imag, axis), axis), axis), axis)
    n = a.shape[-1]
    m = n - 1
    x = np.zeros(a.shape[:-1] + (2 * m,))
    x[..., :m] = a[..., :m]
    y = np.fft.fft(x, axis=-1)[..., :m] * np.exp(-1j * np.pi / (2 * m) * np.arange(m))
    ct = np.zeros(a.shape)
    ct[..., :m] = y.real
    return _back(ct * np.sqrt(2.0 / m), axis)


def xdct3(a, axis):
    """DCT-III on the first N-1 points (last output fixed to zero), norm sqrt(2/(N-1))."""
    a = _mv(np.asarray(a), axis)
    if np.iscomplexobj(a):
        return _back(_mv(xdct3(_back(a.real, axis), axis), axis)
                     + 1j * _mv(xdct3(_back(a.imag, axis), axis), axis), axis)
    n = a.shape[-1]
    m = n - 1
    x = np.zeros(a.shape[:-1] + (2 * m,), dtype=complex)
    x[..., :m] = a[..., :m]
    x[..., 0] = 0.0
    x[..., :m] *= np.exp(-1j * np.pi / (2 * m) * np.arange(m))
    y = np.fft.fft(x, axis=-1)
    ct = np.zeros(a.shape)
    ct[..., :m] = y[..., :m].real + 0.5 * a[..., :1]
    return _back(ct * np.sqrt(2.0 / m), axis)


def xdst1(a, axis):
    """DST-I on interior points (ends set to zero), norm sqrt(2/(N-1)); complex-linear."""
    a = _mv(np.asarray(a), axis)
    n = a.shape[-1]
    ext = np.zeros(a.shape[:-1] + (2 * (n - 1),), dtype=complex)
    ext[..., 1:n - 1] = a[..., 1:n - 1]
    ext[..., n:] = -a[..., n - 2:0:-1]
    y = np.fft.fft(ext, axis=-1)
    out = np.zeros(a.shape, dtype=complex)
    out[..., 1:n - 1] = y[..., 1:n - 1] * np.sqrt(2.0 / (n - 1)) / (-2j)
    if not np.iscomplexobj(a):
        out = out.real
    return _back(out, axis)


def xdst2(a, axis):
    """DST-II on the last N-1 points (first output fixed to zero), norm sqrt(2/(N-1))."""
    a = _mv(np.asarray(a), axis)
    if np.iscomplexobj(a):
        return _back(_mv(xdst2(_back(a.real, axis), axis), axis)
                     + 1j * _mv(xdst2(_back(a.imag, axis), axis), axis), axis)
    n = a.shape[-1]
    m = n - 1
    x = np.zeros(a.shape[:-1] + (2 * m,), dtype=complex)
    x[..., :m] = a[..., 1:n] * np.exp(-1j * np.pi / m * np.arange(m))
    y = np.fft.fft(x, axis=-1)[..., :m] * np.exp(-1j * np.pi / (2 * m) * np.arange(1, m + 1))
    ct = np.zeros(a.shape)
    ct[..., 1:n] = -y.imag
    return _back(ct * np.sqrt(2.0 / m), axis)


def xdst3(a, axis):
    """DST-III on the last N-1 points (first output fixed to zero), norm sqrt(2/(N-1))."""
    a = _mv(np.asarray(a), axis)
    if np.iscomplexobj(a):
        return _back(_mv(xdst3(_back(a.real, axis), axis), axis)
                     + 1j * _mv(xdst3(_back(a.imag, axis), axis), axis), axis)
    n = a.shape[-1]
    m = n - 1
    x = np.zeros(a.shape[:-1] + (2 * m,), dtype=complex)
    x[..., :m] = a[..., 1:n]
    x[..., m - 1] = 0.0
    x[..., :m] *= np.exp(-1j * np.pi / (2 * m) * np.arange(m))
    y = np.fft.fft(x, axis=-1)[..., :m] * np.exp(-1j * np.pi / m * (np.arange(m) + 0.5))
    ct = np.zeros(a.shape)
    alt = np.cumprod(-np.ones(m))
    ct[..., 1:n] = -y.imag - 0.5 * alt * a[..., n - 1:n]
    return _back(ct * np.sqrt(2.0 / m), axis)


# boundary type index  = lower + 2*upper, with -1 Neumann (Robin), +1 Dirichlet,
# 0 periodic (xtransform.m:33-35, spdpreferences.m:53-54)
FORWARD = {-3: xdct1, -1: xdst3, 0: None, 1: xdct3, 3: xdst1}
INVERSE = {-3: xdct1, -1: xdst2, 0: None, 1: xdct2, 3: xdst1}


def transform(a, axis, btype, inverse):
    """One axis of xtransform.m:39-65."""
    if btype not in FORWARD:
        raise ValueError("Invalid boundary type: %r" % (btype,))
    if btype == 0:
        return np.fft.ifft(a, axis=axis) if inverse else np.fft.fft(a, axis=axis)
    f = INVERSE[btype] if inverse else FORWARD[btype]
    return f(a, axis)


# ---------------------------------------------------------------- brute force
def brute_matrix(kind, n):
    """Dense matrix of each transform from its definition (for pinning tests).

    Built from sums written out index by index, independent of the FFT
    restatements above.
    """
    m = n - 1
    T = np.zeros((n, n))
    if kind == "dct1":
        # symmetric extension of length 2m followed by a DFT, first n outputs
        for k in range(n):
            for j in range(n):
                w = 1.0 if j in (0, m) else 2.0
                T[k, j] = w * np.cos(np.pi * j * k / m)
        return T / np.sqrt(2.0 * m)
    if kind == "dst1":
        for k in range(1, m):
            for j in range(1, m):
                T[k, j] = np.sin(np.pi * j * k / m)
        return T * np.sqrt(2.0 / m)
    if kind == "dct2":
        # X_k = sum_{n<m} x_n cos(pi/m (n+1/2) k) * sqrt(2/m), k < m
        for k in range(m):
            for j in range(m):
                T[k, j] = np.cos(np.pi / m * (j + 0.5) * k)
        return T * np.sqrt(2.0 / m)
    if kind == "dct3":
        # X_k = [x_0/2 + sum_{1<=n<m} x_n cos(pi/m (k+1/2) n)] * sqrt(2/m), k < m
        for k in range(m):
            T[k, 0] = 0.5
            for j in range(1, m):
                T[k, j] = np.cos(np.pi / m * (k + 0.5) * j)
        return T * np.sqrt(2.0 / m)
    if kind == "dst2":
        # inputs x_1..x_m (index j=1..m), outputs 1..m
        for k in range(m):
            for j in range(m):
                T[k + 1, j + 1] = np.sin(np.pi / m * (j + 0.5) * (k + 1))
        return T * np.sqrt(2.0 / m)
    if kind == "dst3":
        for k in range(m):
            for j in range(m - 1):
                T[k + 1, j + 1] = np.sin(np.pi / m * (j + 1) * (k + 0.5))
            T[k + 1, m] = 0.5 * (-1.0) ** k
        return T * np.sqrt(2.0 / m)
    raise ValueError(kind)
