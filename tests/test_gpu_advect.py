"""GPU parity of the advected periodic step (pf3d.cpp:399-485: spectral gradient,
force, Stokes projection, N(psi) - U.grad psi) against the FP64 oracle."""
import math
import sys

import numpy as np
import pytest

from helpers import cfg_reference, rel_l2
from oracle import oracle as O
from smectic3d_amd import Solver, SmxError
from smectic3d_amd import _abi as A

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def structured_ic(shape, seed=3):
    """Layers along z with an x/y modulation plus noise, so the force has a solenoidal part."""
    nx, ny, nz = shape
    x = np.arange(nx)[:, None, None] * 2 * math.pi / nx
    y = np.arange(ny)[None, :, None] * 2 * math.pi / ny
    z = np.arange(nz)[None, None, :] * 2 * math.pi * 2 / nz
    psi = 0.8 * np.cos(z + 0.5 * np.sin(x) + 0.3 * np.cos(y)) + O.ic_noise(nx, ny, nz, seed, 0.05)
    return np.ascontiguousarray(psi)


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 16, 24), (24, 32, 32), (256, 16, 16), (384, 16, 16),
                                   (512, 16, 16), (16, 16, 128)])
@pytest.mark.parametrize("force", [A.SMX_FORCE_SH, A.SMX_FORCE_MU])
def test_advect_parity_fp64(shape, force):
    say("advect fp64", shape, force)
    cfg = cfg_reference(*shape, dt=2e-3, advect=1, force_form=force)
    psi = structured_ic(shape)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    with Solver(cfg) as s:
        s.set_psi(psi)
        for n in (1, 9):
            s.step(n)
            orc.step(n)
            e = rel_l2(s.get_psi(), orc.get_psi())
            say("  after", s.step_count(), "rel L2", e)
            assert e < 1e-10
        assert abs(s.energy() - orc.energy()) / abs(orc.energy()) < 1e-12


def test_advect_changes_the_step():
    """Guard that the advected chain is really taken: it differs from the plain step."""
    shape = (32, 16, 24)
    psi = structured_ic(shape)
    outs = []
    for adv in (0, 1):
        cfg = cfg_reference(*shape, dt=2e-3, advect=adv)
        with Solver(cfg) as s:
            s.set_psi(psi)
            s.step(10)
            outs.append(s.get_psi())
    assert rel_l2(outs[1], outs[0]) > 1e-6


def test_advect_penalty_fp64():
    shape = (16, 32, 32)
    say("advect + penalty")
    cfg = cfg_reference(*shape, dt=2e-3, advect=1, w_mid=1.0, mid_plane=16)
    psi = structured_ic(shape)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    orc.step(10)
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(10)
        assert rel_l2(s.get_psi(), orc.get_psi()) < 1e-10


@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_advected_workload_small(prec):
    """Advection-coupled run: eps=0.5, q0=1, gamma=1, dt=0.02, eta=1, mu grad psi force,
    psi0 ~ U[-0.05, 0.05], 8 points per wavelength; 32^3 for 200 steps."""
    n = 32
    L = n / 8 * 2 * math.pi
    mk = lambda p: A.make_config(n, lx=L, ly=L, lz=L, eps=0.5, q0=1.0, beta=0.0, gamma=1.0, dt=0.02,
                                 nl_model=A.SMX_NL_CUBIC, advect=1, eta=1.0, rho=1.0,
                                 force_form=A.SMX_FORCE_MU, precision=p)
    psi = O.ic_noise(n, n, n, 77, 0.05)
    orc = O.Oracle(mk(A.SMX_FP64))
    orc.set_psi(psi)
    orc.step(200)
    with Solver(mk(prec)) as s:
        s.set_psi(psi)
        s.step(200)
        e = rel_l2(s.get_psi(), orc.get_psi())
        say("  advected workload", prec, "rel L2", e)
        assert e < (1e-10 if prec == A.SMX_FP64 else 1e-4)


@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
@pytest.mark.parametrize("force", [A.SMX_FORCE_SH, A.SMX_FORCE_MU])
def test_advect_dtpsi_and_surftrack(prec, force):
    """dpsi/dt of an advected state includes -U.grad psi (surfTrack.cpp:585-606), and the periodic
    surfTrack of the advected loop (surfTrack.cpp is an advected driver) uses it."""
    shape = (32, 16, 48)
    say("advect dtpsi", prec, force)
    cfg = cfg_reference(*shape, dt=2e-3, advect=1, force_form=force, precision=prec)
    psi = structured_ic(shape)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    orc.step(3)
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(3)
        e = rel_l2(s.dtpsi(), orc.dtpsi())
        say("  dtpsi rel L2", e)
        if prec != A.SMX_FP64:  # FP32 states can move a marginal crossing by one cell
            assert e < 1e-5
            return
        (gz, gv, gh, gk), gf = s.surftrack(0.1)
    (oz, ov, oh, ok), of = orc.surftrack(0.1)
    say("  found", gf, of)
    assert e < 1e-10
    assert gf == of and gf > 0 and np.array_equal(gz, oz)
    np.testing.assert_allclose(gv, ov, rtol=0, atol=1e-9 * np.abs(ov).max())


@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_advect_tuned_y_passes(prec):
    """The advected chain runs its eleven y passes with the handle's tuned schedules."""
    shape = (64, 64, 32)
    say("advect autotuned", prec)
    cfg = cfg_reference(*shape, dt=2e-3, advect=1, precision=prec)
    psi = structured_ic(shape)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    orc.step(5)
    with Solver(cfg) as s:
        s.set_schedule([(8, 1), (16, 0), (32, 1), (16, 0)])
        s.set_psi(psi)
        s.step(5)
        e = rel_l2(s.get_psi(), orc.get_psi())
        chosen = s.autotune(3)
        say("  fixed schedule err", e, "autotuned", chosen)
        s.set_psi(psi)
        s.step(5)
        e2 = rel_l2(s.get_psi(), orc.get_psi())
    tol = 1e-10 if prec == A.SMX_FP64 else 2e-5
    assert e < tol and e2 < tol
