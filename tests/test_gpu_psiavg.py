"""psiAvg on the GPU (smx_psi_avg) against the oracle's literal restatement of
cosFFT/comet/morphSurfAmp.cpp:1005-1134, on the solver's own psi: periodic and Neumann,
FP64 and FP32-mixed (the diagnostic runs in FP64 on the state's FP64 image)."""
import sys

import numpy as np
import pytest

from helpers import cfg_reference
from oracle import oracle as O
from smectic3d_amd import Solver, SmxError
from smectic3d_amd import _abi as A

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("bc", [A.SMX_BC_PERIODIC, A.SMX_BC_NEUMANN])
@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_psi_avg_matches_oracle(bc, prec):
    shape, nw = (24, 16, 32), 8
    say("psiAvg", shape, nw, bc, prec)
    cfg = cfg_reference(*shape, eps=0.5, dt=0.05, nl_model=A.SMX_NL_CUBIC, precision=prec, bc=bc)
    with Solver(cfg) as s:
        s.set_psi(O.ic_noise(*shape, 17, 0.3))
        s.step(20)
        psi = s.get_psi()
        got = s.psi_avg(nw)
    want = O.psi_avg(psi, cfg.lx / cfg.nx, cfg.ly / cfg.ny, cfg.lz / cfg.nz, nw, cfg.q0)
    err = np.abs(got - want).max() / np.abs(want).max()
    say("  max rel err", err)
    assert err < 1e-12


def test_psi_avg_rejects_a_window_too_large():
    with Solver(cfg_reference(16, 16, 16)) as s:
        s.set_psi(O.ic_noise(16, 16, 16, 1, 0.1))
        with pytest.raises(SmxError) as ei:
            s.psi_avg(12)  # 3 * 6 > 15
        assert ei.value.code == A.SMX_ERR_INVALID
