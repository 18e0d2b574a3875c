"""Pins of the oracle's transforms (CPU only).

Each transform is checked against (a) a brute-force O(N^2) evaluation of the
definition FFTW documents and (b) an independent library routine (numpy /
scipy pocketfft), so that a dropped term, a wrong sign or a transposed axis
fails at least one of them.
"""
import numpy as np
import pytest
import scipy.fft as sf

from oracle import oracle as O


def brute_dft(x, sign):
    n = len(x)
    j = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(j, j) / n) @ x


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 16, 24, 25, 30, 48, 96, 128, 384])
def test_fft1d_brute_and_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        got = O.fft1d(x, sign)
        np.testing.assert_allclose(got, brute_dft(x, sign), rtol=0, atol=1e-12 * n)
    np.testing.assert_allclose(O.fft1d(x, -1), np.fft.fft(x), rtol=0, atol=1e-13 * n)


def brute_rfft3(a):
    """R2C by the triple-sum definition, kz in [0, nz/2]."""
    nx, ny, nz = a.shape
    out = np.zeros((nx, ny, nz // 2 + 1), complex)
    ex = np.exp(-2j * np.pi * np.outer(np.arange(nx), np.arange(nx)) / nx)
    ey = np.exp(-2j * np.pi * np.outer(np.arange(ny), np.arange(ny)) / ny)
    ez = np.exp(-2j * np.pi * np.outer(np.arange(nz), np.arange(nz // 2 + 1)) / nz)
    out = np.einsum("ijk,ia,jb,kc->abc", a.astype(complex), ex, ey, ez)
    return out


def brute_irfft3_fftw(c, nz):
    """C2R as a multi-dimensional FFTW c2r evaluates it: complex backward DFTs
    along x and y, then a 1-D c2r along z that reads only the real parts of the
    kz = 0 and kz = nz/2 entries (the other half is the Hermitian mirror)."""
    nx, ny, nzh = c.shape
    ex = np.exp(2j * np.pi * np.outer(np.arange(nx), np.arange(nx)) / nx)
    ey = np.exp(2j * np.pi * np.outer(np.arange(ny), np.arange(ny)) / ny)
    t = np.einsum("abc,ai,bj->ijc", c, ex, ey)
    full = np.zeros((nx, ny, nz), complex)
    for k in range(nzh):
        v = t[:, :, k].copy()
        if k == 0 or 2 * k == nz:
            v = v.real + 0j
        full[:, :, k] = v
        if 0 < k and 2 * k != nz:
            full[:, :, nz - k] = np.conj(v)
    ez = np.exp(2j * np.pi * np.outer(np.arange(nz), np.arange(nz)) / nz)
    return np.einsum("ijc,cz->ijz", full, ez).real


@pytest.mark.parametrize("shape", [(4, 6, 8), (3, 5, 8), (6, 6, 12), (5, 4, 6), (2, 3, 4)])
def test_rfft3_brute(shape):
    a = np.random.default_rng(1).standard_normal(shape)
    np.testing.assert_allclose(O.rfft3(a), brute_rfft3(a), rtol=0, atol=1e-12 * a.size)
    np.testing.assert_allclose(O.rfft3(a), np.fft.rfftn(a), rtol=0, atol=1e-12 * a.size)


@pytest.mark.parametrize("shape", [(4, 6, 8), (3, 5, 8), (6, 6, 12), (5, 4, 6)])
def test_irfft3_non_hermitian_semantics(shape):
    """Arbitrary (non-Hermitian) input, as produced by odd-derivative
    multipliers with the +pi/dx Nyquist convention (pf3d.cpp:176,184,404-406)."""
    rng = np.random.default_rng(2)
    nx, ny, nz = shape
    c = rng.standard_normal((nx, ny, nz // 2 + 1)) + 1j * rng.standard_normal((nx, ny, nz // 2 + 1))
    want = brute_irfft3_fftw(c, nz)
    got = O.irfft3(c, nz)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * c.size)
    # numpy/pocketfft irfftn follows the same order and discards the same imaginary parts
    np.testing.assert_allclose(got, np.fft.irfftn(c, s=shape) * np.prod(shape), rtol=0, atol=1e-12 * c.size)


@pytest.mark.parametrize("shape", [(8, 8, 8), (6, 10, 12), (16, 24, 32)])
def test_r2c_c2r_roundtrip_scaling(shape):
    a = np.random.default_rng(3).standard_normal(shape)
    np.testing.assert_allclose(O.irfft3(O.rfft3(a), shape[2]), a * np.prod(shape), rtol=0, atol=1e-13 * np.prod(shape) * np.abs(a).max())


def brute_redft10(x):
    n = len(x)
    j = np.arange(n)
    return 2 * np.cos(np.pi * np.outer(np.arange(n), j + 0.5) / n) @ x


def brute_redft01(x):
    n = len(x)
    j = np.arange(1, n)
    k = np.arange(n)
    return x[0] + 2 * np.cos(np.pi * np.outer(k + 0.5, j) / n) @ x[1:]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 16, 24, 48, 96, 384])
def test_dct_1d_brute_and_scipy(n):
    x = np.random.default_rng(n).standard_normal(n)
    np.testing.assert_allclose(O.redft10_1d(x), brute_redft10(x), rtol=0, atol=1e-12 * n)
    np.testing.assert_allclose(O.redft01_1d(x), brute_redft01(x), rtol=0, atol=1e-12 * n)
    np.testing.assert_allclose(O.redft10_1d(x), sf.dct(x, type=2), rtol=0, atol=1e-12 * n)
    np.testing.assert_allclose(O.redft01_1d(x), sf.dct(x, type=3), rtol=0, atol=1e-12 * n)


@pytest.mark.parametrize("shape", [(4, 6, 8), (3, 5, 7), (8, 8, 12)])
def test_dct_3d(shape):
    a = np.random.default_rng(4).standard_normal(shape)
    np.testing.assert_allclose(O.redft10_3d(a), sf.dctn(a, type=2), rtol=0, atol=1e-11)
    np.testing.assert_allclose(O.redft01_3d(a), sf.dctn(a, type=3), rtol=0, atol=1e-11)
    # REDFT01 o REDFT10 = 2n per axis (cosNoAdvMorph.cpp:240-246)
    np.testing.assert_allclose(O.redft01_3d(O.redft10_3d(a)), a * 8 * np.prod(shape), rtol=1e-12, atol=1e-10)
