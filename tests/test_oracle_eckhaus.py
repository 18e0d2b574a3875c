"""The oracle's Eckhaus scan (energyPsi.m:68-234, oracle/smx_oracle.c or_eckhaus_scan) pinned
by closed forms, and the scan's Q list (energyPsi.m:25,108)."""
import math

import numpy as np

from oracle import oracle as O
from smectic3d_amd import _abi as A
from smectic3d_amd.sweeps import eckhaus_scan_q


def test_constant_field_closed_form():
    """psi0 = c: the z stencil of a constant vanishes, so every entry is the FD + trapz energy
    of c: (-eps/2 c^2 + alpha/2 q0^4 c^2 + f(c)) times (nx-1)dx (ny-1)dy (nz-1)dz (trapz over
    n points weighs the two ends by 1/2)."""
    nx, ny, nz = 10, 12, 14
    cfg = A.make_config(nx, ny, nz, eps=0.4, q0=1.1, beta=0.3, gamma=1.0, nl_model=A.SMX_NL_QUINTIC)
    c = 0.35
    e = O.eckhaus_scan(cfg, np.full((nx, ny, nz), c), [-0.1, 0.0, 0.2, 1.3], [0.0, 0.01, 0.01, 0.5])
    dx, dy, dz = cfg.lx / nx, cfg.ly / ny, cfg.lz / nz
    dens = -0.5 * cfg.eps * c * c + 0.5 * cfg.alpha * (cfg.q0 ** 2 * c) ** 2 - cfg.beta / 4 * c ** 4 + cfg.gamma / 6 * c ** 6
    want = dens * (nx - 1) * dx * (ny - 1) * dy * (nz - 1) * dz
    np.testing.assert_allclose(e, want, rtol=1e-13)


def test_unperturbed_entry_is_the_fd_trapz_energy():
    cfg = A.make_config(16, eps=0.5, beta=0.0, gamma=1.0, nl_model=A.SMX_NL_CUBIC)
    psi = O.ic_noise(16, 16, 16, 4, 0.3)
    o = O.Oracle(cfg)
    o.set_psi(psi)
    e = O.eckhaus_scan(cfg, psi, [-0.2, 0.4], [0.0, 0.0])
    np.testing.assert_allclose(e, o.energy_fd_trapz(), rtol=1e-14)


def test_scan_q_list_is_the_scripts():
    """-dq:dq:1.4 with dq = 2 pi/L, L = np 2 pi/(q0 Nw) for np = 512, Nw = 16: 46 values."""
    L = 512 * 2 * math.pi / 16
    q, amp = eckhaus_scan_q(L, 512)
    dq = 2 * math.pi / L
    assert len(q) == 46 and abs(q[0] + dq) < 1e-15 and abs(q[-1] - 44 * dq) < 1e-12
    assert amp[0] == 0.0 and np.all(amp[1:] == 0.01)
