"""smx_step_noalloc (the stepping entry point that allocates nothing; its z-pass partial sums
live in a static device array) against smx_step, through the C ABI: bit for bit on two calls
in a row, the second on a smaller grid, for the periodic, Neumann and advected steps."""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import cfg_reference
from oracle import oracle as O
from smectic3d_amd import Solver, SmxError
from smectic3d_amd import _abi as A
from smectic3d_amd._lib import lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(cfg, psi, nsteps, noalloc):
    with Solver(cfg) as s:
        s.set_psi(psi)
        fn = lib().smx_step_noalloc if noalloc else lib().smx_step
        rc = fn(s._h, C.c_int64(nsteps))
        assert rc == 0, lib().smx_last_error().decode()
        s.sync()
        return s.get_psi(), s.energy()


CASES = [
    ("periodic fp32", dict(precision=A.SMX_FP32_MIXED), (64, 64, 96), (32, 32, 48)),
    ("periodic fp64", dict(precision=A.SMX_FP64), (48, 64, 32), (32, 16, 16)),
    ("neumann fp32", dict(precision=A.SMX_FP32_MIXED, bc=A.SMX_BC_NEUMANN), (64, 48, 96), (32, 32, 32)),
    ("advected fp32", dict(precision=A.SMX_FP32_MIXED, advect=1, dt=2e-3), (64, 64, 64), (32, 32, 48)),
]


@pytest.mark.parametrize("name,kw,big,small", CASES)
def test_noalloc_step_equals_step(name, kw, big, small):
    say("noalloc", name, big, small)
    for shape, nsteps, seed in ((big, 37, 3), (small, 23, 4)):  # two calls in a row, the second smaller
        cfg = cfg_reference(*shape, **kw)
        psi = O.ic_noise(*shape, seed, 0.05)
        got, eg = run(cfg, psi, nsteps, True)
        want, ew = run(cfg, psi, nsteps, False)
        say("  ", shape, "max diff", float(np.abs(got - want).max()), eg, ew)
        assert np.array_equal(got, want) and eg == ew


def test_noalloc_keeps_no_l1_snapshot():
    cfg = cfg_reference(32, 32, 32, precision=A.SMX_FP32_MIXED)
    with Solver(cfg) as s:
        s.set_psi(O.ic_noise(32, 32, 32, 1, 0.05))
        assert lib().smx_step_noalloc(s._h, C.c_int64(5)) == 0
        with pytest.raises(SmxError) as e:
            s.l1()
        assert e.value.code == A.SMX_ERR_STATE
        assert lib().smx_step_noalloc(s._h, C.c_int64(-1)) == A.SMX_ERR_INVALID
