"""Pins of the oracle's time step, energy and surfTrack against what the
mathematics fixes (CPU only).  None of these restates the oracle's formula:

* linear modes (beta = gamma = 0) evolve by the closed-form CNAB2 factor
  ((1 + dt L/2)/(1 - dt L/2))^n with L(k) = eps - alpha (|k|^2 - q0^2)^2, for
  every mode of a multi-mode field, with k computed independently by numpy;
* the pf3dVal growth law sigma/(1 - dt sigma/2) at lattice wavenumbers;
* second-order convergence of CNAB2 (with its forward-Euler bootstrap) on the
  spatially uniform ODE psi' = (eps - alpha q0^4) psi - gamma psi^3, whose
  exact solution is known;
* single-mode energies from <cos^2> = 1/2, <cos^4> = 3/8, <cos^6> = 5/16;
* decrease of the Lyapunov energy along the gradient flow;
* a z-only profile has a gradient force, which the Stokes projection removes:
  the advected step equals the plain step;
* surfTrack on analytic level sets (planes: H = K = 0; spheres: H = -2/R).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from smectic3d_amd import _abi as A


def kvec(n, L):
    return 2 * np.pi * np.fft.fftfreq(n, d=L / n)


def test_linear_modes_closed_form():
    nx, ny, nz = 12, 16, 10
    cfg = A.make_config(nx, ny, nz, lx=9.0, ly=11.0, lz=7.5, eps=0.4, q0=1.3, alpha=0.8, beta=0.0,
                        gamma=0.0, dt=0.07, nl_model=A.SMX_NL_CUBIC)
    psi = O.ic_noise(nx, ny, nz, 5, 1.0)
    o = O.Oracle(cfg)
    o.set_psi(psi)
    n = 13
    o.step(n)
    kx, ky, kz = kvec(nx, 9.0), kvec(ny, 11.0), kvec(nz, 7.5)
    k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    Lk = 0.4 - 0.8 * (k2 - 1.3 ** 2) ** 2
    g = ((1 + 0.035 * Lk) / (1 - 0.035 * Lk)) ** n
    want = np.real(np.fft.ifftn(np.fft.fftn(psi) * g))
    np.testing.assert_allclose(o.get_psi(), want, rtol=0, atol=1e-12)


def test_growth_law_pf3dval():
    """pf3dVal.cpp:214-326 at lattice wavenumbers mq = 2 pi m / Lz; amplitude 1e-4 makes
    the quintic terms ~1e-8 relative."""
    nz = 32
    cfg = A.make_config(8, 16, nz, nw=16, eps=-0.7, beta=2.0, gamma=1.0, dt=1e-3)
    Lz = cfg.lz
    dmq = 2 * math.pi / Lz
    g = O.growth_sweep(cfg, dmq * 1, dmq, 10)
    mq = dmq * np.arange(1, 11)
    sigma = -0.7 - (mq ** 2 - 1.0) ** 2
    np.testing.assert_allclose(g, sigma / (1 - 1e-3 * sigma / 2), rtol=1e-6)


def uniform_run(dt, T, psi0=0.3, eps=0.5, gamma=1.0, q0=1.0):
    cfg = A.make_config(4, 16, 4, eps=eps, q0=q0, beta=0.0, gamma=gamma, dt=dt, nl_model=A.SMX_NL_CUBIC)
    o = O.Oracle(cfg)
    o.set_psi(np.full((4, 16, 4), psi0))
    o.step(int(round(T / dt)))
    return o.get_psi()[0, 0, 0]


def test_cnab2_second_order_on_uniform_ode():
    eps, q0, gamma, p0, T = 0.5, 1.0, 1.0, 0.3, 2.0
    a = eps - q0 ** 4
    exact = math.copysign(math.sqrt(a / (gamma + (a / p0 ** 2 - gamma) * math.exp(-2 * a * T))), p0)
    errs = [abs(uniform_run(dt, T, p0, eps, gamma, q0) - exact) for dt in (0.04, 0.02, 0.01)]
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.6 < r1 < 4.4 and 3.6 < r2 < 4.4, (errs, r1, r2)


def test_bootstrap_first_step_uniform():
    """Step 1 uses N_past = N(psi0) (pf3d.cpp:363-374), so for a uniform field it is
    psi1 = ((1 + dt L0/2) c + dt N(c)) / (1 - dt L0/2), L0 = eps - alpha q0^4."""
    c, eps, q0, beta, gamma, dt = 0.7, -0.7, 1.0, 2.0, 1.0, 0.01
    cfg = A.make_config(4, 16, 4, eps=eps, q0=q0, beta=beta, gamma=gamma, dt=dt)
    o = O.Oracle(cfg)
    o.set_psi(np.full((4, 16, 4), c))
    o.step(1)
    L0 = eps - q0 ** 4
    N = beta * c ** 3 - gamma * c ** 5
    assert abs(o.get_psi()[1, 3, 2] - ((1 + dt * L0 / 2) * c + dt * N) / (1 - dt * L0 / 2)) < 1e-15


@pytest.mark.parametrize("model", [A.SMX_NL_CUBIC, A.SMX_NL_QUINTIC])
def test_single_mode_energy(model):
    nx, ny, nz = 8, 16, 24
    lx, ly, lz = 10.0, 12.0, 14.0
    eps, q0, alpha, beta, gamma, amp = 0.3, 1.1, 0.9, 0.4, 1.2, 0.8
    cfg = A.make_config(nx, ny, nz, lx=lx, ly=ly, lz=lz, eps=eps, q0=q0, alpha=alpha, beta=beta,
                        gamma=gamma, nl_model=model)
    q = 2 * math.pi * 3 / lz
    z = np.arange(nz) * lz / nz
    psi = np.broadcast_to(amp * np.cos(q * z), (nx, ny, nz)).copy()
    o = O.Oracle(cfg)
    o.set_psi(psi)
    V = lx * ly * lz
    quad = (alpha / 2 * (q0 ** 2 - q ** 2) ** 2 - eps / 2) * amp ** 2 / 2
    if model == A.SMX_NL_CUBIC:  # <cos^3> = 0, <cos^4> = 3/8
        loc = gamma / 4 * amp ** 4 * 3 / 8
    else:  # -beta/4 <cos^4> + gamma/6 <cos^6>
        loc = -beta / 4 * amp ** 4 * 3 / 8 + gamma / 6 * amp ** 6 * 5 / 16
    assert abs(o.energy() - V * (quad + loc)) < 1e-12 * V


def test_energy_fd_trapz_closed_form():
    """energyPsi.m:217-234 on psi = A cos(q z): the periodic 3-point Laplacian has the exact
    eigenvalue -(4/dz^2) sin^2(q dz/2), and trapz over n points (end weights 1/2) of
    cos^2 / cos^4 on a lattice mode is n/2 - (1 + cos^2(q dz))/2 and 3n/8 - (1 + cos^4(q dz))/2."""
    nx, ny, nz = 6, 8, 32
    lx, ly, lz = 5.0, 7.0, 2 * math.pi * 4
    eps, q0, alpha, gamma, amp = 0.3, 1.0, 1.0, 1.0, 0.5
    cfg = A.make_config(nx, ny, nz, lx=lx, ly=ly, lz=lz, eps=eps, q0=q0, alpha=alpha, beta=0.0, gamma=gamma,
                        nl_model=A.SMX_NL_CUBIC)
    dz = lz / nz
    q = 2 * math.pi * 3 / lz
    z = np.arange(nz) * dz
    psi = np.broadcast_to(amp * np.cos(q * z), (nx, ny, nz)).copy()
    o = O.Oracle(cfg)
    o.set_psi(psi)
    lam = -4 / dz ** 2 * math.sin(q * dz / 2) ** 2
    c2 = nz / 2 - (1 + math.cos(q * dz) ** 2) / 2
    c4 = 3 * nz / 8 - (1 + math.cos(q * dz) ** 4) / 2
    per_z = ((-eps / 2 + alpha / 2 * (q0 ** 2 + lam) ** 2) * amp ** 2 * c2 + gamma / 4 * amp ** 4 * c4) * dz
    want = (nx - 1) * (lx / nx) * (ny - 1) * (ly / ny) * per_z
    assert abs(o.energy_fd_trapz() - want) < 1e-12 * abs(want)


def test_energy_decreases():
    cfg = A.make_config(16, 16, 16, nw=8, eps=0.5, beta=0.2, gamma=1.0, dt=0.02, nl_model=A.SMX_NL_CUBIC)
    o = O.Oracle(cfg)
    o.set_psi(O.ic_noise(16, 16, 16, 9, 0.3))
    es = []
    for _ in range(8):
        o.step(25)
        es.append(o.energy())
    assert all(b <= a + 1e-12 for a, b in zip(es, es[1:])), es


def test_stokes_removes_gradient_force():
    """A psi that depends on z only makes S = psiSH dpsi/dz e_z a gradient; its
    projection (pf3d.cpp:433-447) is zero, so the advected step equals the plain one."""
    n = 16
    z = np.arange(n) * (2 * math.pi * 2 / n)
    prof = 0.6 * np.cos(z) + 0.2 * np.sin(2 * z) + 0.1 * np.cos(3 * z)
    psi = np.broadcast_to(prof, (8, 16, n)).copy()
    outs = []
    for adv in (0, 1):
        cfg = A.make_config(8, 16, n, lx=8.0, ly=9.0, lz=2 * math.pi * 2, eps=-0.7, beta=2.0,
                            gamma=1.0, dt=1e-3, advect=adv)
        o = O.Oracle(cfg)
        o.set_psi(psi)
        o.step(20)
        outs.append(o.get_psi())
    np.testing.assert_allclose(outs[1], outs[0], rtol=0, atol=1e-13)


def test_stokes_advects_non_gradient_force():
    """A field varying in x and z makes a force with a solenoidal part: advection then acts."""
    nx, ny, nz = 16, 16, 16
    x = np.arange(nx) * (2 * math.pi * 2 / nx)
    z = np.arange(nz) * (2 * math.pi * 2 / nz)
    psi = (0.6 * np.cos(z)[None, None, :] * (1 + 0.3 * np.cos(x)[:, None, None])
           + 0.2 * np.sin(x)[:, None, None] * np.sin(2 * z)[None, None, :])
    psi = np.broadcast_to(psi, (nx, ny, nz)).copy()
    outs = []
    for adv in (0, 1):
        cfg = A.make_config(nx, ny, nz, lx=2 * math.pi * 2, ly=9.0, lz=2 * math.pi * 2, eps=-0.7, beta=2.0,
                            gamma=1.0, dt=1e-2, advect=adv)
        o = O.Oracle(cfg)
        o.set_psi(psi)
        o.step(10)
        outs.append(o.get_psi())
    assert np.abs(outs[1] - outs[0]).max() > 1e-8


def test_surftrack_neumann_plane_and_sphere():
    nx, ny, nz = 24, 24, 48
    L = 12.0
    dx = L / nx
    cfg = A.make_config(nx, ny, nz, lx=L, ly=L, lz=2 * L, bc=A.SMX_BC_NEUMANN, eps=0.3, beta=0.0,
                        gamma=1.0, nl_model=A.SMX_NL_CUBIC)
    x = (np.arange(nx) + 0.5) * dx
    z = (np.arange(nz) + 0.5) * dx
    # plane: psi > 0 above z0, < 0 below (scan from the top arms, then crosses at z0)
    z0 = 10.3
    psi = np.broadcast_to(np.tanh((z - z0) / 1.5), (nx, ny, nz)).copy()
    o = O.Oracle(cfg)
    o.set_psi(psi)
    (sz, v, H, K), found = o.surftrack(0.7)
    assert found == nx * ny
    assert np.all(np.abs(sz - round(z0 / dx - 0.5)) <= 1)
    assert np.abs(H).max() < 1e-10 and np.abs(K).max() < 1e-10


def _sphere_errors(nx, R=4.0, w=1.5, L=12.0):
    nz = 2 * nx
    dx = L / nx
    cfg = A.make_config(nx, nx, nz, lx=L, ly=L, lz=2 * L, bc=A.SMX_BC_NEUMANN, eps=0.3, beta=0.0,
                        gamma=1.0, nl_model=A.SMX_NL_CUBIC)
    x = (np.arange(nx) + 0.5) * dx
    z = (np.arange(nz) + 0.5) * dx
    r = np.sqrt((x[:, None, None] - L / 2) ** 2 + (x[None, :, None] - L / 2) ** 2 + (z[None, None, :] - L) ** 2)
    o = O.Oracle(cfg)
    o.set_psi(np.tanh((R - r) / w))  # psi > 0 inside the sphere
    prev = [np.full(nx * nx, np.nan) for _ in range(4)]
    (sz, v, H, K), found = o.surftrack(0.7, prev=prev)
    hit = np.isfinite(sz)
    assert found == hit.sum() and found > 0.4 * math.pi * (R / dx) ** 2
    xs = x - L / 2
    rho = np.sqrt(xs[:, None] ** 2 + xs[None, :] ** 2).ravel()
    # the level set through the selected grid point is a sphere of radius rr:
    # H = div(grad psi/|grad psi|) = -2/rr, K = 1/rr^2
    rr = np.sqrt(rho ** 2 + ((sz + 0.5) * dx - L) ** 2)
    eh = np.abs(H[hit] * rr[hit] / -2 - 1).max()
    ek = np.abs(K[hit] * rr[hit] ** 2 - 1).max()
    return eh, ek


def test_surftrack_neumann_sphere_second_order():
    """Megrabov H and K (cosNoAdvMorphCurv.cpp:945-961) on spheres converge to -2/r and
    1/r^2 at second order in dx, over every column that crosses (normals from vertical to
    oblique), so a dropped or mis-signed mixed-derivative term fails."""
    eh1, ek1 = _sphere_errors(32)
    eh2, ek2 = _sphere_errors(64)
    assert eh1 < 0.03 and ek1 < 0.06 and eh2 < 0.008 and ek2 < 0.016
    assert 3.0 < eh1 / eh2 < 5.5 and 3.0 < ek1 / ek2 < 5.5


def test_surftrack_periodic_plane():
    nx, ny, nz = 16, 16, 32
    cfg = A.make_config(nx, ny, nz, lx=8.0, ly=8.0, lz=16.0, eps=0.3, beta=0.0, gamma=1.0,
                        nl_model=A.SMX_NL_CUBIC)
    z = np.arange(nz) * 0.5
    psi = np.broadcast_to(np.cos(2 * math.pi * z / 16.0), (nx, ny, nz)).copy()
    o = O.Oracle(cfg)
    o.set_psi(psi)
    (sz, v, H, K), found = o.surftrack(0.1)
    assert found == nx * ny
    # cos(pi k/16) is +6e-17 at k = 8 and first negative at k = 9; the scan moves to the
    # smaller |psi| (k = 8) for the derivatives but reports k = 9 (surfTrack.cpp:779-785)
    assert np.all(sz == 9)
    assert np.abs(H).max() < 1e-12 and np.abs(K).max() < 1e-12


def test_splitmix_known_answer():
    # first SplitMix64 output for seed 0 is 0xE220A8397B1DCDAF
    assert O.splitmix_uniform(0, 0, 0) == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53


def test_surftrack_periodic_curved_surface_closed_form():
    """Periodic surfTrack on psi = cos(q (z - h(x, y))), h = a cos(2 pi x/Lx) cos(2 pi y/Ly): the
    crossing layer z = h + pi/(2q) is curved.  At the reported cell the oracle's
    H = (psi_xx + psi_yy)/|grad psi| and K = (psi_xx psi_yy - psi_xy^2)/|grad psi|^2
    (surfTrack.cpp:612-617,789-797) must equal the same expressions of the analytic derivatives
    there (the field is band-limited to 1e-20, so spectral derivatives are exact to rounding)."""
    nx, ny, nz = 32, 24, 48
    Lx, Ly, Lz = 10.0, 8.0, 12.0
    q = 2 * math.pi * 2 / Lz
    a = 0.25
    x = np.arange(nx) * Lx / nx
    y = np.arange(ny) * Ly / ny
    z = np.arange(nz) * Lz / nz
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    kx, ky = 2 * math.pi / Lx, 2 * math.pi / Ly
    h = a * np.cos(kx * X) * np.cos(ky * Y)
    hx = -a * kx * np.sin(kx * X) * np.cos(ky * Y)
    hy = -a * ky * np.cos(kx * X) * np.sin(ky * Y)
    hxx, hyy = -kx * kx * h, -ky * ky * h
    hxy = a * kx * ky * np.sin(kx * X) * np.sin(ky * Y)
    th = q * (Z - h)
    psi = np.cos(th)
    s, c = np.sin(th), np.cos(th)
    pxx = q * hxx * s - q * q * hx * hx * c
    pyy = q * hyy * s - q * q * hy * hy * c
    pxy = q * hxy * s - q * q * hx * hy * c
    g = np.sqrt((q * hx * s) ** 2 + (q * hy * s) ** 2 + (q * s) ** 2)
    cfg = A.make_config(nx, ny, nz, lx=Lx, ly=Ly, lz=Lz, eps=0.3, beta=0.0, gamma=1.0, nl_model=A.SMX_NL_CUBIC)
    o = O.Oracle(cfg)
    o.set_psi(psi)
    (sz, v, H, K), found = o.surftrack(0.1)
    assert found == nx * ny
    Hs = H.reshape(nx, ny)
    Ks = K.reshape(nx, ny)
    checked = 0
    for i in range(nx):
        for j in range(ny):
            k = int(sz[i * ny + j])
            kk = k - 1 if abs(psi[i, j, k]) > abs(psi[i, j, k - 1]) else k  # the cell the reference evaluates
            He = (pxx[i, j, kk] + pyy[i, j, kk]) / g[i, j, kk]
            Ke = (pxx[i, j, kk] * pyy[i, j, kk] - pxy[i, j, kk] ** 2) / g[i, j, kk] ** 2
            assert abs(Hs[i, j] - He) < 1e-10 * (1 + abs(He))
            assert abs(Ks[i, j] - Ke) < 1e-10 * (1 + abs(Ke))
            checked += 1
    assert checked == nx * ny and np.ptp(Hs) > 1e-2  # the surface is curved
