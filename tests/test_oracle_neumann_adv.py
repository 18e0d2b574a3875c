"""Oracle checks of the advected Neumann step (sine/cosine Stokes with per-component parity).

The reference builds this machinery (morphSurfK.cpp:331-348,392-418,864-882) but never
applies it (:896-898), so nothing of the reference pins it directly.  It is pinned instead by
the periodic advected step (pf3d.cpp:399-465, itself pinned in test_oracle_step.py): a
Neumann box is the periodic box mirrored evenly to [0, 2L) along every axis.
  * One step from any state equals one periodic step of the mirrored state, symmetrised over
    the 8 mirror images.  The symmetrisation removes what the periodic step's Nyquist
    convention (+pi/dx for odd derivatives, pf3d.cpp:176,184) adds at the top mode: with
    k_x = +pi/dx only, the projection couples the x-Nyquist force into U_y and U_z with one
    sign, which breaks the mirror symmetry; a Neumann box has no such term.
  * For fields without top-mode content the periodic step keeps the mirror symmetry itself,
    and many steps agree without symmetrising.
The 1-D DST-II/III are checked against their defining sums and scipy.fft (pocketfft), whose
unnormalised type-2/3 transforms follow FFTW's RODFT10/RODFT01 definitions.
"""
import math

import numpy as np
import pytest
import scipy.fft as sf

from oracle import oracle as O
from smectic3d_amd import _abi as A


def _mirror(psi):
    e = np.concatenate([psi, psi[::-1]], 0)
    e = np.concatenate([e, e[:, ::-1]], 1)
    return np.concatenate([e, e[:, :, ::-1]], 2)


def _cfgs(shape, L, **kw):
    base = dict(eps=-0.7, beta=2.0, gamma=1.0, dt=2e-2, advect=1, force_form=A.SMX_FORCE_SH)
    base.update(kw)
    nx, ny, nz = shape
    cn = A.make_config(nx, ny, nz, lx=L[0], ly=L[1], lz=L[2], bc=A.SMX_BC_NEUMANN, **base)
    cp = A.make_config(2 * nx, 2 * ny, 2 * nz, lx=2 * L[0], ly=2 * L[1], lz=2 * L[2], **base)
    return cn, cp


def _smooth(shape, L):
    nx, ny, nz = shape
    x = (np.arange(nx) + 0.5) * L[0] / nx
    y = (np.arange(ny) + 0.5) * L[1] / ny
    z = (np.arange(nz) + 0.5) * L[2] / nz
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    return 0.5 * np.cos(4 * np.pi * Z / L[2]) * (1 + 0.4 * np.cos(3 * np.pi * Y / L[1]) * np.cos(np.pi * X / L[0]))


@pytest.mark.parametrize("n", [1, 2, 5, 8, 12])
def test_dst_definitions(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    j = np.arange(n)
    d2 = np.array([2 * np.sum(x * np.sin(np.pi * (j + 0.5) * (k + 1) / n)) for k in range(n)])
    d3 = np.array([(-1) ** k * x[n - 1] + 2 * np.sum(x[: n - 1] * np.sin(np.pi * (j[: n - 1] + 1) * (k + 0.5) / n))
                   for k in range(n)])
    np.testing.assert_allclose(O.rodft10_1d(x), d2, rtol=0, atol=1e-12)
    np.testing.assert_allclose(O.rodft01_1d(x), d3, rtol=0, atol=1e-12)
    np.testing.assert_allclose(O.rodft10_1d(x), sf.dst(x, type=2), rtol=0, atol=1e-12)
    np.testing.assert_allclose(O.rodft01_1d(x), sf.dst(x, type=3), rtol=0, atol=1e-12)
    np.testing.assert_allclose(O.rodft01_1d(O.rodft10_1d(x)), 2 * n * x, rtol=0, atol=1e-11 * n)


@pytest.mark.parametrize("form", [A.SMX_FORCE_SH, A.SMX_FORCE_MU])
@pytest.mark.parametrize("shape,L", [((8, 6, 10), (7.0, 5.5, 9.0)), ((6, 8, 8), (6.0, 8.0, 7.5))])
def test_one_step_is_symmetrised_mirror_step(form, shape, L):
    """Full-spectrum noise (top modes included): Neumann step == mirror-symmetrised periodic step."""
    rng = np.random.default_rng(7)
    psi = 0.3 * rng.standard_normal(shape)
    cn, cp = _cfgs(shape, L, force_form=form)
    on = O.Oracle(cn)
    on.set_psi(psi)
    on.step(1)
    op = O.Oracle(cp)
    op.set_psi(_mirror(psi))
    op.step(1)
    b = op.get_psi()
    s = (b + b[::-1]) / 2
    s = (s + s[:, ::-1]) / 2
    s = (s + s[:, :, ::-1]) / 2
    nx, ny, nz = shape
    a = on.get_psi()
    np.testing.assert_allclose(a, s[:nx, :ny, :nz], rtol=0, atol=1e-13)
    # the periodic step alone is not mirror symmetric for such data (its +pi/dx Nyquist)
    assert np.abs(a - b[:nx, :ny, :nz]).max() > 1e-6


@pytest.mark.parametrize("form", [A.SMX_FORCE_SH, A.SMX_FORCE_MU])
def test_many_steps_equal_mirror_for_smooth_fields(form):
    shape, L = (8, 8, 8), (7.0, 6.0, 9.0)
    psi = _smooth(shape, L)
    cn, cp = _cfgs(shape, L, force_form=form)
    on = O.Oracle(cn)
    on.set_psi(psi)
    on.step(20)
    op = O.Oracle(cp)
    op.set_psi(_mirror(psi))
    op.step(20)
    np.testing.assert_allclose(on.get_psi(), op.get_psi()[:8, :8, :8], rtol=0, atol=1e-13)
    # advection does act on this field
    c0 = A.make_config(8, 8, 8, lx=L[0], ly=L[1], lz=L[2], bc=A.SMX_BC_NEUMANN, eps=-0.7, beta=2.0, gamma=1.0,
                       dt=2e-2, advect=0)
    o0 = O.Oracle(c0)
    o0.set_psi(psi)
    o0.step(20)
    assert np.abs(o0.get_psi() - on.get_psi()).max() > 1e-6


def test_gradient_force_is_not_advected():
    """psi(z) only: S = f d_z psi e_z is a gradient, its projection vanishes."""
    nz = 16
    z = (np.arange(nz) + 0.5) * (9.0 / nz)
    prof = 0.6 * np.cos(2 * np.pi * z / 4.5) + 0.2 * np.cos(3 * np.pi * z / 9.0)
    psi = np.broadcast_to(prof, (8, 8, nz)).copy()
    outs = []
    for adv in (0, 1):
        c = A.make_config(8, 8, nz, lx=6.0, ly=7.0, lz=9.0, bc=A.SMX_BC_NEUMANN, eps=-0.7, beta=2.0, gamma=1.0,
                          dt=1e-2, advect=adv)
        o = O.Oracle(c)
        o.set_psi(psi)
        o.step(10)
        outs.append(o.get_psi())
    np.testing.assert_allclose(outs[1], outs[0], rtol=0, atol=1e-13)


def test_velocity_is_divergence_free_stokes_flow():
    """Below the top modes, U solves rho eta |K|^2 U = S - K p (one scalar p per K) with K.U = 0,
    checked with scipy's DCT/DST spectra of U and of the force built here from psi."""
    shape, L = (8, 6, 10), (7.0, 5.5, 9.0)
    rng = np.random.default_rng(3)
    psi = 0.3 * rng.standard_normal(shape)
    eta, rho = 1.7, 1.3
    cn, _ = _cfgs(shape, L, eta=eta, rho=rho)
    o = O.Oracle(cn)
    o.set_psi(psi)
    U = o.velocity()
    n = shape
    q = [np.pi * np.arange(m + 1) / l for m, l in zip(n, L)]  # wavenumbers pi K / L, K = 0..n
    ph = sf.dctn(psi, type=2)
    qq = [q[a][:n[a]] for a in range(3)]
    k2 = qq[0][:, None, None] ** 2 + qq[1][None, :, None] ** 2 + qq[2][None, None, :] ** 2
    sh = sf.idctn(cn.alpha * (k2 - cn.q0 ** 2) ** 2 * ph, type=2)  # psiSH
    spec = []
    for a in range(3):
        c = np.zeros_like(ph)
        dst = [slice(None)] * 3
        src = [slice(None)] * 3
        dst[a], src[a] = slice(0, n[a] - 1), slice(1, n[a])
        w = [1, 1, 1]
        w[a] = n[a] - 1
        c[tuple(dst)] = -q[a][1:n[a]].reshape(w) * ph[tuple(src)]
        g = c
        for b in range(3):
            g = sf.idst(g, type=2, axis=b) if b == a else sf.idct(g, type=2, axis=b)
        S = sh * g
        Sh, Uh = S, U[a]
        for b in range(3):
            Sh = sf.dst(Sh, type=2, axis=b) if b == a else sf.dct(Sh, type=2, axis=b)
            Uh = sf.dst(Uh, type=2, axis=b) if b == a else sf.dct(Uh, type=2, axis=b)
        spec.append((Sh, Uh))
    # wavevectors with every K_a in 1..n_a - 1 hold all three components
    sl = lambda a, b: slice(0, n[b] - 2) if b == a else slice(1, n[b] - 1)
    Kq = np.meshgrid(*[q[b][1:n[b] - 1] for b in range(3)], indexing="ij")
    Sx = [spec[a][0][tuple(sl(a, b) for b in range(3))] for a in range(3)]
    Ux = [spec[a][1][tuple(sl(a, b) for b in range(3))] for a in range(3)]
    K2 = Kq[0] ** 2 + Kq[1] ** 2 + Kq[2] ** 2
    div = sum(Kq[a] * Ux[a] for a in range(3))
    scale = max(np.abs(u).max() for u in Ux) * math.sqrt(K2.max())
    assert np.abs(div).max() < 1e-12 * scale
    R = [Sx[a] - rho * eta * K2 * Ux[a] for a in range(3)]  # must be K p
    for a in range(3):
        for b in range(a + 1, 3):
            cross = R[a] * Kq[b] - R[b] * Kq[a]
            assert np.abs(cross).max() < 1e-11 * max(np.abs(s).max() for s in Sx) * K2.max()
    assert max(np.abs(u).max() for u in Ux) > 1e-3


@pytest.mark.parametrize("advect", [0, 1])
@pytest.mark.parametrize("bc", [A.SMX_BC_PERIODIC, A.SMX_BC_NEUMANN])
def test_dtpsi_is_the_step_rate(bc, advect):
    """dpsi/dt = F^-1[s (L psi_hat + N_hat)] with the advective term in N (surfTrack.cpp:585-606):
    after a first step, (psi^2 - psi^1)/dt = [L psi^1 + N^1 + (N^1 - N^0)/2] / C2 approaches it at
    first order in dt.  The advective term is 5-10% of dpsi/dt here, so leaving it out would stall
    the convergence at that level."""
    shape = (16, 16, 24)
    x = (np.arange(16) + 0.5) / 16 * 2 * np.pi
    z = (np.arange(24) + 0.5) / 24 * 2 * np.pi
    psi = (0.7 * np.cos(2 * z)[None, None, :] * (1 + 0.3 * np.cos(x)[:, None, None])
           + 0.05 * np.sin(x)[:, None, None] * np.cos(3 * z)[None, None, :])
    psi = np.broadcast_to(psi, shape).copy()
    errs = []
    for dt in (4e-5, 2e-5, 1e-5):
        cfg = A.make_config(*shape, lx=8.0, ly=7.0, lz=11.0, bc=bc, eps=0.3, beta=0.0, gamma=1.0,
                            nl_model=A.SMX_NL_CUBIC, dt=dt, advect=advect, force_form=A.SMX_FORCE_MU)
        o = O.Oracle(cfg)
        o.set_psi(psi)
        o.step(1)
        p1 = o.get_psi()
        d = o.dtpsi()
        o.step(1)
        errs.append(np.linalg.norm((o.get_psi() - p1) / dt - d) / np.linalg.norm(d))
    assert errs[-1] < 1e-2
    for a, b in zip(errs, errs[1:]):
        assert 1.8 < a / b < 2.2, errs
