"""Error of the FP32-mixed path against the FP64 path on the 256^3 workload (hyperboloid IC
+ 1e-3 noise, L = 32*2pi, eps = 0.3, dt = 0.1, cubic model) over its full 2000 steps.

The targets of 1e-6 relative L2 on psi and 1e-9 relative energy are not met by an FP32 state:
every stored value carries a relative rounding of ~6e-8, and on this pattern-forming trajectory
a per-step perturbation of that size is amplified by ~1e4 over 2000 steps (the FP64 path's own
drift against the oracle grows from 1e-15 to 1e-10 the same way).  This test states the levels
the mixed path does reach, as regression bounds: psi within 2e-3, energy within 5e-7 and the
dominant wavenumber within 2e-7 after 2000 steps, and psi within 5e-6 after 25 steps.  (The
energy error depends on the rounding path: 4.5e-9 to 1.1e-7 across tile schedules.)"""
import math
import sys

import numpy as np
import pytest

from helpers import cfg_hyperboloid_256, rel_l2
from oracle import oracle as O
from smectic3d_amd import Solver
from smectic3d_amd import _abi as A

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_mixed_path_error_on_the_256_workload():
    n = 256
    dx = 2 * math.pi / 8
    psi0 = O.ic_hyperboloid(n, n, n, dx, dx, dx, 1.0, 0.4, 1e-3, 0)
    out = {}
    for prec in (A.SMX_FP64, A.SMX_FP32_MIXED):
        with Solver(cfg_hyperboloid_256(prec)) as s:
            s.set_psi(psi0)
            s.step(25)
            early = s.get_psi()
            s.step(1975)
            out[prec] = (early, s.get_psi(), s.energy(), s.dominant_wavenumber())
    e64, p64, E64, q64 = out[A.SMX_FP64]
    e32, p32, E32, q32 = out[A.SMX_FP32_MIXED]
    err = {"psi_25": rel_l2(e32, e64), "psi_2000": rel_l2(p32, p64), "energy": abs(E32 - E64) / abs(E64),
           "qdom": abs(q32 - q64) / abs(q64)}
    say("  FP32-mixed vs FP64:", err)
    assert err["psi_25"] < 5e-6
    assert err["psi_2000"] < 2e-3
    assert err["energy"] < 5e-7
    assert err["qdom"] < 2e-7
