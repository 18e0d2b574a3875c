"""GPU parity of the Neumann (DCT-II/III) step, cosFFT/mesabi/cosNoAdvMorph.cpp:483-530,
against the FP64 oracle.  Tolerances as for the periodic path."""
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


def cfg_n(shape, precision=A.SMX_FP64, **kw):
    base = dict(bc=A.SMX_BC_NEUMANN, eps=0.5, beta=0.3, gamma=1.0, dt=0.05, nl_model=A.SMX_NL_CUBIC, nw=8)
    base.update(kw)
    return A.make_config(*shape, precision=precision, **base)


LENGTHS = [16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512]


def _axis_shapes():
    out = []
    for n in LENGTHS:
        out.append((n, 16, 16))
        if n % 16 == 0:
            out.append((16, n, 16))
        out.append((16, 16, n))
    return out


@pytest.mark.parametrize("shape", _axis_shapes())
@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_neumann_every_length(shape, prec):
    say("neumann length", shape, prec)
    cfg = cfg_n(shape, prec)
    psi = O.ic_noise(*shape, 23, 0.3)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    orc.step(5)
    with Solver(cfg) as s:
        s.set_psi(psi)
        r0 = rel_l2(s.get_psi(), psi)
        s.step(5)
        e = rel_l2(s.get_psi(), orc.get_psi())
    say("  roundtrip", r0, "5-step", e)
    tol = 1e-11 if prec == A.SMX_FP64 else 2e-5
    assert r0 < tol and e < tol


@pytest.mark.parametrize("model", [A.SMX_NL_QUINTIC, A.SMX_NL_CUBIC])
def test_neumann_parity_fp64(model):
    shape = (32, 48, 64)
    say("neumann parity", shape, model)
    if model == A.SMX_NL_QUINTIC:
        cfg = cfg_n(shape, eps=-0.7, beta=2.0, dt=5e-3, nl_model=model, nw=16)
        amp = 0.8
    else:
        cfg = cfg_n(shape)
        amp = 0.2
    psi = O.ic_noise(*shape, 31, amp)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    with Solver(cfg) as s:
        s.set_psi(psi)
        for n in (1, 10, 60):
            s.step(n)
            orc.step(n)
            e = rel_l2(s.get_psi(), orc.get_psi())
            say("  after", s.step_count(), "rel L2", e)
            assert e < 1e-10
        assert abs(s.l1() - orc.l1()) / orc.l1() < 1e-9
        ge, oe = s.energy(), orc.energy()
        say("  energy", ge, oe)
        assert abs(ge - oe) / abs(oe) < 1e-12
        assert rel_l2(s.dtpsi(), orc.dtpsi()) < 1e-10
        gq, oq = s.dominant_wavenumber(), orc.dominant_wavenumber()
        assert abs(gq - oq) / oq < 1e-12


def test_neumann_hyperboloid_cfg_small():
    """The Neumann workload (eps=0.5, gamma=1, dt=0.05, cubic, 8 points per wavelength,
    hyperboloid IC) at 32x32x48 for 400 steps, FP64 and FP32-mixed against the oracle."""
    import bench
    nx, ny, nz = 32, 32, 48
    dx = 2 * math.pi / 8
    cfg64 = cfg_n((nx, ny, nz), eps=0.5, beta=0.0, gamma=1.0, dt=0.05)
    psi = O.ic_hyperboloid(nx, ny, nz, dx, dx, dx, 1.0, 0.4, 1e-3, 0)
    orc = O.Oracle(cfg64)
    orc.set_psi(psi)
    orc.step(400)
    with Solver(cfg64) as s:
        s.set_psi(psi)
        s.step(400)
        e = rel_l2(s.get_psi(), orc.get_psi())
        ge, oe = s.energy(), orc.energy()
        say("  fp64 rel L2", e, "energy", abs(ge - oe) / abs(oe))
        assert e < 1e-10 and abs(ge - oe) / abs(oe) < 1e-12
    cfg32 = cfg_n((nx, ny, nz), A.SMX_FP32_MIXED, eps=0.5, beta=0.0, gamma=1.0, dt=0.05)
    with Solver(cfg32) as s:
        s.set_psi(psi)
        s.step(400)
        e = rel_l2(s.get_psi(), orc.get_psi())
        say("  fp32-mixed rel L2", e)
        assert e < 1e-4
