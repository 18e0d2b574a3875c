"""The Eckhaus scan on the GPU (smx_eckhaus_scan) against the oracle's restatement of
energyPsi.m:68-234 on the solver's own state, FP64 and FP32-mixed (the scan runs in FP64)."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from smectic3d_amd import Solver
from smectic3d_amd import _abi as A
from smectic3d_amd.sweeps import eckhaus_scan, eckhaus_scan_q

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_eckhaus_scan_matches_oracle(prec):
    shape = (24, 16, 32)
    print("eckhaus scan", shape, prec, flush=True)
    sys.stdout.flush()
    cfg = A.make_config(*shape, eps=-0.7, beta=2.0, gamma=1.0, dt=1e-3, nl_model=A.SMX_NL_QUINTIC, precision=prec)
    with Solver(cfg) as s:
        s.set_psi(O.ic_noise(*shape, 11, 0.4))
        s.step(30)
        psi0 = s.get_psi()
        q, e = eckhaus_scan(s, qmax=1.4)
    _, amp = eckhaus_scan_q(cfg.lz, cfg.nz, 1.4)
    want = O.eckhaus_scan(cfg, psi0, q, amp)
    err = np.abs(e - want).max() / np.abs(want).max()
    print("  entries", len(q), "max rel err", err, flush=True)
    assert len(q) == len(want) and err < 1e-12
