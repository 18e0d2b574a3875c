"""Cross-check of the periodic step against torch.fft on the same GPU (rocFFT underneath):
an independent FFT library, used here only as a checker.  The CNAB2 step of pf3d.cpp:395-485
without advection is restated with torch's rfftn/irfftn in FP64 (tests/torch_step_ref.py), with
the first N_past from N(psi0) (pf3d.cpp:363-374).  torch's own HIP runtime does not start in a
process where libsmectic's already runs, so the restatement runs in a child process and writes
its result to a file."""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import rel_l2
from oracle import oracle as O
from smectic3d_amd import Solver
from smectic3d_amd import _abi as A

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def torch_steps(cfg, psi0, nsteps, tmp_path):
    import importlib.util
    if importlib.util.find_spec("torch") is None:
        pytest.skip("torch is not installed")
    src, dst = tmp_path / "in.npz", tmp_path / "out.npz"
    np.savez(src, psi0=psi0, nsteps=nsteps, lx=cfg.lx, ly=cfg.ly, lz=cfg.lz, eps=cfg.eps, alpha=cfg.alpha,
             q0=cfg.q0, beta=cfg.beta, gamma=cfg.gamma, dt=cfg.dt, cubic=float(cfg.nl_model == A.SMX_NL_CUBIC))
    r = subprocess.run([sys.executable, os.path.join(HERE, "torch_step_ref.py"), str(src), str(dst)],
                       capture_output=True, text=True, timeout=240)
    say(r.stdout.strip()[-300:], r.stderr.strip()[-300:])
    assert r.returncode == 0, r.stderr[-2000:]
    return np.load(dst)["psi"]


@pytest.mark.parametrize("shape,model", [((32, 32, 48), A.SMX_NL_CUBIC), ((64, 48, 32), A.SMX_NL_QUINTIC),
                                         ((96, 64, 128), A.SMX_NL_CUBIC)])
def test_periodic_step_matches_torch_fft(shape, model, tmp_path):
    say("torch.fft cross-check", shape, model)
    nx, ny, nz = shape
    cfg = A.make_config(nx, ny, nz, nw=8, eps=0.5, beta=0.3 if model == A.SMX_NL_CUBIC else 2.0, gamma=1.0,
                        dt=0.02, nl_model=model)
    psi0 = O.ic_noise(nx, ny, nz, 11, 0.3)
    ref = torch_steps(cfg, psi0, 10, tmp_path)
    with Solver(cfg) as s:
        s.set_psi(psi0)
        s.step(10)
        got = s.get_psi()
    e = rel_l2(got, ref)
    say("  rel L2 vs torch.fft", e)
    assert e < 1e-11
