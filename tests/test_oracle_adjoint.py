"""The oracle's discrete adjoint (FP64 reference for the GPU adjoint) against
central finite differences of the forward oracle: dJ/deps, dJ/dq0 and the
directional derivative of dJ/dpsi0 (CPU only)."""
import numpy as np
import pytest

from oracle import oracle as O
from smectic3d_amd import _abi as A

BASE = dict(nw=8, eps=0.4, q0=1.0, beta=0.3, gamma=1.0, dt=0.05, nl_model=A.SMX_NL_CUBIC)


def cfg_with(shape, ref, **kw):
    c = A.make_config(*shape, **{**BASE, **kw})
    c.lx, c.ly, c.lz = ref.lx, ref.ly, ref.lz  # the grid does not move with q0
    return c


@pytest.mark.parametrize("model", [A.SMX_NL_CUBIC, A.SMX_NL_QUINTIC])
def test_adjoint_matches_central_differences(model):
    shape, T, h = (16, 16, 16), 12, 1e-6
    kw = dict(nl_model=model) if model == A.SMX_NL_CUBIC else dict(nl_model=model, eps=-0.2, beta=1.0, dt=0.01)
    cfg = cfg_with(shape, A.make_config(*shape, nw=8), **kw)
    psi = O.ic_noise(*shape, 3, 0.3 if model == A.SMX_NL_CUBIC else 0.6)
    J, d, de, dq = O.adjoint_grad(cfg, psi, T)
    assert abs(J - O.loss_only(cfg, psi, T)) < 1e-12 * J
    e0, q0 = cfg.eps, cfg.q0
    fe = (O.loss_only(cfg_with(shape, cfg, **{**kw, "eps": e0 + h}), psi, T)
          - O.loss_only(cfg_with(shape, cfg, **{**kw, "eps": e0 - h}), psi, T)) / (2 * h)
    fq = (O.loss_only(cfg_with(shape, cfg, **{**kw, "q0": q0 + h}), psi, T)
          - O.loss_only(cfg_with(shape, cfg, **{**kw, "q0": q0 - h}), psi, T)) / (2 * h)
    assert abs(de - fe) <= 1e-7 * max(1.0, abs(fe))
    assert abs(dq - fq) <= 1e-7 * max(1.0, abs(fq))
    v = O.ic_noise(*shape, 9, 1.0)
    fd = (O.loss_only(cfg, psi + h * v, T) - O.loss_only(cfg, psi - h * v, T)) / (2 * h)
    assert abs(np.sum(d * v) - fd) <= 1e-7 * max(1.0, abs(fd))


def test_adjoint_single_step_is_transpose():
    """T = 1 with beta = gamma = 0: dJ/dpsi0 = A^T (psi_1 - g) with A = F^-1 (C1/C2) F,
    checked by the exact closed form through numpy's FFT."""
    shape = (8, 16, 12)
    cfg = A.make_config(*shape, nw=8, eps=0.3, q0=1.1, beta=0.0, gamma=0.0, dt=0.1, nl_model=A.SMX_NL_CUBIC)
    psi = O.ic_noise(*shape, 4, 1.0)
    J, d, de, dq = O.adjoint_grad(cfg, psi, 1, amp=0.5)
    k = [2 * np.pi * np.fft.fftfreq(n, d=L / n) for n, L in zip(shape, (cfg.lx, cfg.ly, cfg.lz))]
    k2 = k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2 + k[2][None, None, :] ** 2
    Lk = 0.3 - (k2 - 1.1 ** 2) ** 2
    g = (1 + 0.05 * Lk) / (1 - 0.05 * Lk)
    psi1 = np.real(np.fft.ifftn(np.fft.fftn(psi) * g))
    z = np.arange(shape[2]) * cfg.lz / shape[2]
    a1 = psi1 - 0.5 * np.cos(1.1 * z)
    np.testing.assert_allclose(d, np.real(np.fft.ifftn(np.fft.fftn(a1) * g)), rtol=0, atol=1e-12)
