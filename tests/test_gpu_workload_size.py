"""Step parity at the benchmarked workloads' own grids, under the autotuned tile schedules that
the timed runs install: the FP64 GPU step against the FP64 oracle after a few steps of

  * the periodic pf3dMorph run, 256^3 (hyperboloid start + 1e-3 noise, eps 0.3, dt 0.1, cubic),
  * the advected periodic run, 256^3 (mu grad psi force, eta 1, eps 0.5, dt 0.02),
  * the cosNoAdvMorph Neumann run, 256x256x384 (eps 0.5, dt 0.05, cubic),
  * the advected Neumann run, 256x256x384 (sine/cosine Stokes projection, dt 0.02).

The small-grid tests cover every transform length and tile variant; these cover the large
lengths (256, 384) in the tuned launch shapes, whose grids hold 8 to 40 times more work groups.
FP64 tolerances: relative L2(psi) and L2(d psi/dt) <= 1e-11, relative energy <= 1e-12.  The
FP32-mixed path run over the same steps stays within 1e-5 of the FP64 oracle.  The oracle takes
about 0.5 s (plain), 3 s (advected), 2.5 s (Neumann) and 11 s (advected Neumann) per step on
8 host threads, which sets the step counts."""
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


DX = 2 * math.pi / 8  # 8 points per wavelength of q0 = 1


def periodic(prec):
    L = 256 * DX
    return A.make_config(256, lx=L, ly=L, lz=L, eps=0.3, q0=1.0, beta=0.0, gamma=1.0, dt=0.1,
                         nl_model=A.SMX_NL_CUBIC, precision=prec)


def advected(prec):
    L = 256 * DX
    return A.make_config(256, lx=L, ly=L, lz=L, eps=0.5, q0=1.0, beta=0.0, gamma=1.0, dt=0.02,
                         nl_model=A.SMX_NL_CUBIC, advect=1, eta=1.0, rho=1.0, force_form=A.SMX_FORCE_MU,
                         precision=prec)


def neumann(prec):
    nx, ny, nz = 256, 256, 384
    return A.make_config(nx, ny, nz, lx=nx * DX, ly=ny * DX, lz=nz * DX, bc=A.SMX_BC_NEUMANN, eps=0.5,
                         q0=1.0, beta=0.0, gamma=1.0, dt=0.05, nl_model=A.SMX_NL_CUBIC, precision=prec)


def neumann_advected(prec):
    nx, ny, nz = 256, 256, 384
    return A.make_config(nx, ny, nz, lx=nx * DX, ly=ny * DX, lz=nz * DX, bc=A.SMX_BC_NEUMANN, eps=0.5,
                         q0=1.0, beta=0.0, gamma=1.0, dt=0.02, nl_model=A.SMX_NL_CUBIC, advect=1, eta=1.0,
                         rho=1.0, force_form=A.SMX_FORCE_MU, precision=prec)


def layered(shape):
    nx, ny, nz = shape
    return O.ic_hyperboloid(nx, ny, nz, DX, DX, DX, 1.0, 0.4, 1e-3, 0)


def small_noise(shape):
    return np.random.default_rng(5).uniform(-0.05, 0.05, shape)


# (name, config builder, initial condition, steps)
CASES = [
    ("periodic", periodic, layered, 6),
    ("advected", advected, small_noise, 2),
    ("neumann", neumann, layered, 3),
    ("neumann_advected", neumann_advected, small_noise, 2),
]


@pytest.mark.parametrize("name,mk,ic,steps", CASES, ids=[c[0] for c in CASES])
def test_step_parity_at_workload_size(name, mk, ic, steps):
    cfg = mk(A.SMX_FP64)
    shape = (cfg.nx, cfg.ny, cfg.nz)
    say("workload-size parity", name, shape, steps, "steps")
    psi0 = ic(shape)
    orc = O.Oracle(cfg)
    orc.set_psi(psi0)
    orc.step(steps)
    opsi, oE, odt = orc.get_psi(), orc.energy(), orc.dtpsi()
    del orc
    with Solver(cfg) as s:
        s.set_psi(psi0)
        sched = s.autotune(5)
        say("  tuned schedule", sched)
        s.step(steps)
        e, eE, edt = rel_l2(s.get_psi(), opsi), abs(s.energy() - oE) / abs(oE), rel_l2(s.dtpsi(), odt)
    say("  fp64: rel L2 psi", e, "energy", eE, "dtpsi", edt)
    assert e < 1e-11 and eE < 1e-12 and edt < 1e-11
    with Solver(mk(A.SMX_FP32_MIXED)) as s:
        s.set_psi(psi0)
        s.autotune(5)
        s.step(steps)
        e32 = rel_l2(s.get_psi(), opsi)
    say("  fp32-mixed: rel L2 psi", e32)
    assert e32 < 1e-5
