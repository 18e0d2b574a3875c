"""GPU adjoint of the periodic step (smx_adjoint_grad) against the FP64 oracle's
discrete adjoint and against central finite differences of the GPU forward."""
import copy
import math
import sys

import numpy as np
import pytest

from oracle import oracle as O
from smectic3d_amd import Solver, SmxError
from smectic3d_amd import _abi as A

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def base_cfg(shape, prec=A.SMX_FP64, **kw):
    d = dict(nw=8, eps=0.4, q0=1.0, beta=0.3, gamma=1.0, dt=0.05, nl_model=A.SMX_NL_CUBIC)
    d.update(kw)
    return A.make_config(*shape, precision=prec, **d)


@pytest.mark.parametrize("shape,T,K", [((16, 16, 16), 12, 5), ((32, 16, 24), 20, 20), ((24, 32, 32), 9, 4)])
def test_adjoint_matches_oracle(shape, T, K):
    say("adjoint vs oracle", shape, T, K)
    cfg = base_cfg(shape)
    psi = O.ic_noise(*shape, 3, 0.3)
    J0, d0, e0, q0 = O.adjoint_grad(cfg, psi, T)
    with Solver(cfg) as s:
        J, d, e, q = s.adjoint_grad(psi, T, K)
    say("  J", J, J0, "deps", e, e0, "dq0", q, q0)
    assert abs(J - J0) < 1e-12 * abs(J0)
    assert abs(e - e0) < 1e-9 * max(1, abs(e0)) and abs(q - q0) < 1e-9 * max(1, abs(q0))
    assert np.linalg.norm(d - d0) < 1e-9 * np.linalg.norm(d0)


def test_adjoint_quintic_matches_oracle():
    shape, T = (16, 16, 32), 15
    cfg = base_cfg(shape, eps=-0.2, beta=1.0, dt=0.01, nl_model=A.SMX_NL_QUINTIC)
    psi = O.ic_noise(*shape, 5, 0.6)
    J0, d0, e0, q0 = O.adjoint_grad(cfg, psi, T)
    with Solver(cfg) as s:
        J, d, e, q = s.adjoint_grad(psi, T, 4)
    assert abs(e - e0) < 1e-9 * max(1, abs(e0)) and abs(q - q0) < 1e-9 * max(1, abs(q0))
    assert np.linalg.norm(d - d0) < 1e-9 * np.linalg.norm(d0)


def gpu_loss(cfg, psi, T, amp=0.5):
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(T)
        p = s.get_psi()
    z = np.arange(cfg.nz) * cfg.lz / cfg.nz
    return 0.5 * float(np.sum((p - amp * np.cos(cfg.q0 * z)) ** 2))


def test_adjoint_central_differences_64():
    shape, T, h = (64, 64, 64), 100, 1e-5
    say("adjoint FD 64^3 x 100")
    cfg = base_cfg(shape, beta=0.0)
    psi = O.ic_noise(*shape, 7, 0.05)
    with Solver(cfg) as s:
        J, d, de, dq = s.adjoint_grad(psi, T, 25)

    def with_(**kw):
        c = copy.copy(cfg)
        for k, v in kw.items():
            setattr(c, k, v)
        return c
    fe = (gpu_loss(with_(eps=cfg.eps + h), psi, T) - gpu_loss(with_(eps=cfg.eps - h), psi, T)) / (2 * h)
    fq = (gpu_loss(with_(q0=cfg.q0 + h), psi, T) - gpu_loss(with_(q0=cfg.q0 - h), psi, T)) / (2 * h)
    v = O.ic_noise(*shape, 11, 1.0)
    fv = (gpu_loss(cfg, psi + h * v, T) - gpu_loss(cfg, psi - h * v, T)) / (2 * h)
    say("  deps", de, fe, "dq0", dq, fq, "dpsi.v", float(np.sum(d * v)), fv)
    assert abs(de - fe) < 1e-6 * max(1, abs(fe))
    assert abs(dq - fq) < 1e-6 * max(1, abs(fq))
    assert abs(np.sum(d * v) - fv) < 1e-6 * max(1, abs(fv))


def test_adjoint_central_differences_256_2000():
    """The single large case: 256^3, 2000 steps, checkpoints every 50 (L = 32 wavelengths,
    eps = 0.3, q0 = 1, gamma = 1, dt = 0.05, psi0 noise 0.05, loss 1/2 |psi_T - 0.5 cos(q0 z)|^2),
    FP64, against central differences of the forward solve in eps, q0 and a random psi0
    direction, Richardson-extrapolated from steps h and 2h (the 2000-step map amplifies
    perturbations, so the h^2 term of a plain central difference is visible at 1e-6)."""
    n, T, K = 256, 2000, 50
    say("adjoint FD 256^3 x 2000")
    cfg = A.make_config(n, lx=n / 8 * 2 * math.pi, ly=n / 8 * 2 * math.pi, lz=n / 8 * 2 * math.pi, eps=0.3,
                        q0=1.0, beta=0.0, gamma=1.0, dt=0.05, nl_model=A.SMX_NL_CUBIC)
    psi = np.random.default_rng(9).uniform(-0.05, 0.05, (n, n, n))
    with Solver(cfg) as s:
        J, d, de, dq = s.adjoint_grad(psi, T, K)
        res, rec, _ = s.adjoint_info()
    say("  J", J, "resident", res, "recomputed", rec)

    def with_(**kw):
        c = copy.copy(cfg)
        for k, v in kw.items():
            setattr(c, k, v)
        return c
    h = 1e-6
    v = np.random.default_rng(10).uniform(-1, 1, (n, n, n))

    def central(f, step):
        return (f(step) - f(-step)) / (2 * step)

    def richardson(f):
        d1, d2 = central(f, h), central(f, 2 * h)
        return (4 * d1 - d2) / 3, d1

    fe, fe1 = richardson(lambda t: gpu_loss(with_(eps=cfg.eps + t), psi, T))
    fq, fq1 = richardson(lambda t: gpu_loss(with_(q0=cfg.q0 + t), psi, T))
    fv, fv1 = richardson(lambda t: gpu_loss(cfg, psi + t * v, T))
    dv = float(np.sum(d * v))
    say("  deps", de, fe, fe1, "dq0", dq, fq, fq1, "dpsi.v", dv, fv, fv1)
    assert abs(de - fe) < 1e-6 * abs(fe)
    assert abs(dq - fq) < 1e-6 * abs(fq)
    assert abs(dv - fv) < 1e-6 * abs(fv)


def test_adjoint_fp32_mixed_close_to_fp64():
    shape, T = (32, 32, 32), 40
    psi = O.ic_noise(*shape, 3, 0.1)
    with Solver(base_cfg(shape)) as s:
        J, d, e, q = s.adjoint_grad(psi, T, 10)
    with Solver(base_cfg(shape, A.SMX_FP32_MIXED)) as s:
        J2, d2, e2, q2 = s.adjoint_grad(psi, T, 10)
    say("  fp32 adjoint rel err", abs(e2 - e) / abs(e), abs(q2 - q) / abs(q), np.linalg.norm(d2 - d) / np.linalg.norm(d))
    assert abs(e2 - e) < 1e-4 * abs(e) and abs(q2 - q) < 1e-4 * abs(q)
    assert np.linalg.norm(d2 - d) < 1e-4 * np.linalg.norm(d)


def test_autograd_function():
    torch = pytest.importorskip("torch")
    from smectic3d_amd.autograd import SHLoss
    shape, T = (16, 16, 16), 10
    cfg = base_cfg(shape)
    psi = torch.tensor(O.ic_noise(*shape, 3, 0.3), requires_grad=True)
    eps = torch.tensor(0.4, dtype=torch.float64, requires_grad=True)
    q0 = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    J = SHLoss.apply(psi, eps, q0, cfg, T, 4, 0.5)
    J.backward()
    J0, d0, e0, q00 = O.adjoint_grad(cfg, psi.detach().numpy(), T)
    assert abs(J.item() - J0) < 1e-12 * J0
    assert abs(eps.grad.item() - e0) < 1e-9 * max(1, abs(e0))
    assert abs(q0.grad.item() - q00) < 1e-9 * max(1, abs(q00))
    assert np.linalg.norm(psi.grad.numpy() - d0) < 1e-9 * np.linalg.norm(d0)


def test_adjoint_rejects_unsupported():
    cfg = base_cfg((16, 16, 16), advect=1)
    with Solver(cfg) as s:
        with pytest.raises(SmxError):
            s.adjoint_grad(O.ic_noise(16, 16, 16, 1, 0.1), 5, 2)


@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_adjoint_memory_budget_does_not_change_the_gradient(prec):
    """Resident record sets (forward sweep records, no recompute) against checkpoints only
    (every segment recomputed once): the same kernels on the same states, so the
    gradient is identical; and a partial budget in between."""
    shape, T, K = (32, 32, 48), 23, 5
    say("adjoint budget", prec)
    cfg = base_cfg(shape, prec)
    psi = O.ic_noise(*shape, 13, 0.2)
    out = {}
    with Solver(cfg) as s:
        n = 32 * 32 * 48
        set_bytes = (4 if prec else 8) * n * K + (8 if prec else 16) * 32 * 32 * 25 * (K + 1)
        for name, budget in (("auto", -1), ("none", 0), ("two", 2 * set_bytes)):
            s.set_adjoint_budget(budget)
            out[name] = s.adjoint_grad(psi, T, K)
            res, rec, nbytes = s.adjoint_info()
            say("  ", name, "resident", res, "recomputed", rec, "bytes", nbytes)
            assert res + rec == T
            if name == "auto":
                assert rec == 0
            elif name == "none":
                assert res == 0
            else:  # the last two segments (3 + K steps) resident
                assert res == 3 + K
        back = s.get_psi()  # the handle is left at psi0 (up to the transform round trip)
        assert np.linalg.norm(back - psi) < (1e-14 if prec == A.SMX_FP64 else 2e-6) * np.linalg.norm(psi)
    for name in ("none", "two"):
        J, d, e, q = out[name]
        J0, d0, e0, q0 = out["auto"]
        assert J == J0 and e == e0 and q == q0
        np.testing.assert_array_equal(d, d0)
