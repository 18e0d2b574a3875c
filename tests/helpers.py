"""Shared test helpers: named configurations and error measures."""
import math

import numpy as np

from smectic3d_amd import _abi as A


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cfg_plain_64(precision=A.SMX_FP64, n=64):
    """Plain periodic SH: 64^3, L = 8*2pi/q0, q0=1, eps=0.5, beta=0, gamma=1, dt=0.05, cubic."""
    L = 8 * 2 * math.pi * n / 64
    return A.make_config(n, lx=L, ly=L, lz=L, eps=0.5, q0=1.0, beta=0.0, gamma=1.0, dt=0.05,
                         nl_model=A.SMX_NL_CUBIC, precision=precision)


def cfg_hyperboloid_256(precision=A.SMX_FP64, n=256):
    """pf3dMorph-style periodic run: L = 32*2pi (8 points per wavelength), eps=0.3, dt=0.1, cubic."""
    L = n / 8 * 2 * math.pi
    return A.make_config(n, lx=L, ly=L, lz=L, eps=0.3, q0=1.0, beta=0.0, gamma=1.0, dt=0.1,
                         nl_model=A.SMX_NL_CUBIC, precision=precision)


def cfg_reference(nx, ny=None, nz=None, precision=A.SMX_FP64, **kw):
    """The reference's own constants (pf3d.cpp:46-75: eps=-0.7, beta=2, quintic, Nw=16)."""
    base = dict(eps=-0.7, beta=2.0, gamma=1.0, q0=1.0, dt=1e-3, nl_model=A.SMX_NL_QUINTIC)
    base.update(kw)
    return A.make_config(nx, ny, nz, precision=precision, **base)
