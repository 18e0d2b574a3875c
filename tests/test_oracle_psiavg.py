"""The oracle's psiAvg (cosFFT/comet/morphSurfAmp.cpp:1005-1134, restated literally in
oracle/smx_oracle.c or_psi_avg) pinned by closed forms independent of its loops."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_constant_field_closed_form():
    """psi = c: every window integral is c^2 l1 l2 l3 and the reference box gives
    c^2 (2h)^3 dx dy dz, so psiAvg = lambda^3 / ((2h)^3 dx dy dz) everywhere."""
    nx, ny, nz, nw = 13, 15, 17, 8
    dx, dy, dz, q0 = 0.6, 0.7, 0.8, 1.3
    out = O.psi_avg(np.full((nx, ny, nz), -0.4), dx, dy, dz, nw, q0)
    want = (2 * math.pi / q0) ** 3 / ((nw // 2 * 2) ** 3 * dx * dy * dz)
    np.testing.assert_allclose(out, want, rtol=1e-13)


def _trap_window(f, h, d):
    """Trapezoid integrals of f over [c-h, c+h] clipped to the array, by numpy."""
    n = len(f)
    out = np.zeros(n)
    for c in range(n):
        a, b = max(c - h, 0), min(c + h, n - 1)
        out[c] = np.trapezoid(f[a:b + 1], dx=d) if b > a else 0.0
    return out


def test_separable_field_matches_product_of_1d_trapezoids():
    """psi = a(x) b(y) c(z): the 3-D trapezoid sum factorises, so psiAvg is a product of 1-D
    window integrals (numpy) over the same products on the reference box, times lambda^3 over
    the clipped window volume."""
    nx, ny, nz, nw = 14, 16, 19, 6
    h = nw // 2
    dx, dy, dz = 0.5, 0.45, 0.55
    rng = np.random.default_rng(3)
    a, b, c = rng.uniform(-1, 1, nx), rng.uniform(-1, 1, ny), rng.uniform(-1, 1, nz)
    psi = a[:, None, None] * b[None, :, None] * c[None, None, :]
    out = O.psi_avg(psi, dx, dy, dz, nw)
    tx, ty, tz = _trap_window(a * a, h, dx), _trap_window(b * b, h, dy), _trap_window(c * c, h, dz)
    ref = tx[2 * h] * ty[2 * h] * tz[2 * h]
    lx = np.array([(min(i + h, nx - 1) - max(i - h, 0)) * dx for i in range(nx)])
    ly = np.array([(min(j + h, ny - 1) - max(j - h, 0)) * dy for j in range(ny)])
    lz = np.array([(min(k + h, nz - 1) - max(k - h, 0)) * dz for k in range(nz)])
    want = (tx[:, None, None] * ty[None, :, None] * tz[None, None, :]) * (2 * math.pi) ** 3 / (
        ref * lx[:, None, None] * ly[None, :, None] * lz[None, None, :])
    np.testing.assert_allclose(out, want, rtol=1e-12)


def test_reference_box_must_fit():
    with pytest.raises(ValueError):
        O.psi_avg(np.ones((8, 8, 8)), 1.0, 1.0, 1.0, 6)  # 3 * 3 > 7
