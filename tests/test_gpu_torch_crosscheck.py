"""Cross-check of the periodic step against torch.fft on the same GPU (rocFFT underneath):
an independent FFT library, used here only as a checker.  The CNAB2 step of pf3d.cpp:395-485
without advection is restated with torch's rfftn/irfftn in FP64, with the first N_past from
N(psi0) (pf3d.cpp:363-374)."""
import math
import sys

import numpy as np
import pytest

from helpers import rel_l2
from oracle import oracle as O
from smectic3d_amd import Solver
from smectic3d_amd import _abi as A

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def torch_gpu():
    """torch with its own HIP runtime on the GPU, or a skip: that runtime does not start in a
    process where libsmectic's has already run (the whole GPU suite), so the check runs on its
    own (pytest tests/test_gpu_torch_crosscheck.py)."""
    torch = pytest.importorskip("torch")
    try:
        torch.zeros(1, device="cuda:0")
    except RuntimeError as e:
        pytest.skip(f"torch cannot use the GPU in this process ({e})")
    return torch


def torch_steps(cfg, psi0, nsteps):
    torch = torch_gpu()
    dev = torch.device("cuda:0")
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    kx = 2 * math.pi * torch.fft.fftfreq(nx, d=cfg.lx / nx, dtype=torch.float64, device=dev)
    ky = 2 * math.pi * torch.fft.fftfreq(ny, d=cfg.ly / ny, dtype=torch.float64, device=dev)
    kz = 2 * math.pi * torch.fft.rfftfreq(nz, d=cfg.lz / nz, dtype=torch.float64, device=dev)
    k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    lin = cfg.eps - cfg.alpha * (k2 - cfg.q0 ** 2) ** 2
    c1, c2 = 1 + 0.5 * cfg.dt * lin, 1 - 0.5 * cfg.dt * lin

    def nl(p):
        if cfg.nl_model == A.SMX_NL_CUBIC:
            return cfg.beta * p * p - cfg.gamma * p ** 3
        return cfg.beta * p ** 3 - cfg.gamma * p ** 5

    psi = torch.tensor(psi0, dtype=torch.float64, device=dev)
    npast = torch.fft.rfftn(nl(psi))
    for _ in range(nsteps):
        ph = torch.fft.rfftn(psi)
        nq = torch.fft.rfftn(nl(psi))
        ph = (c1 * ph + 0.5 * cfg.dt * (3 * nq - npast)) / c2
        npast = nq
        psi = torch.fft.irfftn(ph, s=(nx, ny, nz))
    return psi.cpu().numpy()


@pytest.mark.parametrize("shape,model", [((32, 32, 48), A.SMX_NL_CUBIC), ((64, 48, 32), A.SMX_NL_QUINTIC),
                                         ((96, 64, 128), A.SMX_NL_CUBIC)])
def test_periodic_step_matches_torch_fft(shape, model):
    say("torch.fft cross-check", shape, model)
    nx, ny, nz = shape
    cfg = A.make_config(nx, ny, nz, nw=8, eps=0.5, beta=0.3 if model == A.SMX_NL_CUBIC else 2.0, gamma=1.0,
                        dt=0.02, nl_model=model)
    psi0 = O.ic_noise(nx, ny, nz, 11, 0.3)
    ref = torch_steps(cfg, psi0, 10)
    with Solver(cfg) as s:
        s.set_psi(psi0)
        s.step(10)
        got = s.get_psi()
    e = rel_l2(got, ref)
    say("  rel L2 vs torch.fft", e)
    assert e < 1e-11
