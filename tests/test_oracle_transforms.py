"""Pins oracle/transforms.py: FFT restatements of xdct1..xdst3 against brute-force
sums, scipy's textbook transforms, and the inverse pairs xtransform.m relies on."""
import numpy as np
import pytest
import scipy.fft as sf

from oracle import transforms as T

SIZES = [5, 8, 9, 16, 33]
rng = np.random.default_rng(1234)


def _vec(n, cplx=True):
    x = rng.standard_normal((3, n))
    return x + 1j * rng.standard_normal((3, n)) if cplx else x


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("kind", ["dct1", "dct2", "dct3", "dst1", "dst2", "dst3"])
def test_against_brute_force(kind, n):
    f = getattr(T, "x" + kind)
    M = T.brute_matrix(kind, n)
    x = _vec(n)
    np.testing.assert_allclose(f(x, 1), x @ M.T, rtol=1e-12, atol=1e-12)
    # axis handling: transform along axis 0 of a transposed array
    np.testing.assert_allclose(f(x.T, 0), (x @ M.T).T, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n", SIZES)
def test_scipy_types(n):
    x = _vec(n, cplx=False)
    m = n - 1
    # DCT-I: unnormalised scipy type 1 / sqrt(2(N-1))
    np.testing.assert_allclose(T.xdct1(x, 1), sf.dct(x, type=1, axis=1) / np.sqrt(2 * m), atol=1e-12)
    # DST-I on the interior points
    ref = np.zeros_like(x)
    ref[:, 1:m] = sf.dst(x[:, 1:m], type=1, axis=1) / 2 * np.sqrt(2.0 / m)
    np.testing.assert_allclose(T.xdst1(x, 1), ref, atol=1e-12)
    # DCT-II on the first N-1 points
    ref = np.zeros_like(x)
    ref[:, :m] = sf.dct(x[:, :m], type=2, axis=1) / 2 * np.sqrt(2.0 / m)
    np.testing.assert_allclose(T.xdct2(x, 1), ref, atol=1e-12)
    # DST-II on the last N-1 points
    ref = np.zeros_like(x)
    ref[:, 1:] = sf.dst(x[:, 1:], type=2, axis=1) / 2 * np.sqrt(2.0 / m)
    np.testing.assert_allclose(T.xdst2(x, 1), ref, atol=1e-12)


@pytest.mark.parametrize("n", SIZES)
def test_inverse_pairs(n):
    m = n - 1
    x = _vec(n)
    np.testing.assert_allclose(T.xdct1(T.xdct1(x, 1), 1), x, atol=1e-12)        # self-inverse
    xi = x.copy()
    xi[:, 0] = xi[:, -1] = 0
    np.testing.assert_allclose(T.xdst1(T.xdst1(x, 1), 1), xi, atol=1e-12)       # self-inverse (interior)
    # ND: forward xdct3, inverse xdct2 on the first N-1 points (last fixed 0)
    y = x.copy()
    y[:, m] = 0
    np.testing.assert_allclose(T.xdct2(T.xdct3(y, 1), 1), y, atol=1e-12)
    # DN: forward xdst3, inverse xdst2 on the last N-1 points (first fixed 0)
    y = x.copy()
    y[:, 0] = 0
    np.testing.assert_allclose(T.xdst2(T.xdst3(y, 1), 1), y, atol=1e-12)


def test_invalid_boundary_raises():
    with pytest.raises(ValueError):
        T.transform(np.zeros((1, 8)), 1, 2, False)
