"""GPU parity of the advected Neumann step (DCT grid of cosNoAdvMorph.cpp with the
sine/cosine Stokes projection of per-component parity) against the FP64 oracle, whose
formulation is pinned by the mirrored periodic step (test_oracle_neumann_adv.py)."""
import math
import sys

import numpy as np
import pytest

from helpers import rel_l2
from oracle import oracle as O
from smectic3d_amd import Solver
from smectic3d_amd import _abi as A

pytestmark = pytest.mark.gpu

LENGTHS = [16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cfg_na(shape, precision=A.SMX_FP64, **kw):
    base = dict(bc=A.SMX_BC_NEUMANN, eps=-0.7, beta=2.0, gamma=1.0, dt=2e-3, nl_model=A.SMX_NL_QUINTIC, advect=1,
                force_form=A.SMX_FORCE_SH, nw=8)
    base.update(kw)
    return A.make_config(*shape, precision=precision, **base)


def structured_ic(shape, seed=5):
    """Layers along z bent in x and y plus noise: the force has a solenoidal part, and the noise
    puts content in every slot including the top ones."""
    nx, ny, nz = shape
    x = (np.arange(nx)[:, None, None] + 0.5) * math.pi / nx
    y = (np.arange(ny)[None, :, None] + 0.5) * math.pi / ny
    z = (np.arange(nz)[None, None, :] + 0.5) * math.pi * 3 / nz
    psi = 0.8 * np.cos(z + 0.6 * np.cos(x) + 0.3 * np.cos(2 * y)) + O.ic_noise(nx, ny, nz, seed, 0.05)
    return np.ascontiguousarray(psi)


def _axis_shapes():
    out = []
    for n in LENGTHS:
        out.append((n, 16, 16))
        if n % 16 == 0:
            out.append((16, n, 16))
        out.append((16, 16, n))
    return sorted(set(out))


@pytest.mark.parametrize("shape", _axis_shapes())
@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_neumann_adv_every_length(shape, prec):
    say("neumann advected length", shape, prec)
    cfg = cfg_na(shape, prec)
    psi = structured_ic(shape)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    orc.step(5)
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(5)
        e = rel_l2(s.get_psi(), orc.get_psi())
    say("  5-step rel L2", e)
    assert e < (1e-11 if prec == A.SMX_FP64 else 2e-5)


@pytest.mark.parametrize("shape", [(32, 16, 24), (24, 32, 48), (48, 48, 32), (64, 32, 64)])
@pytest.mark.parametrize("force", [A.SMX_FORCE_SH, A.SMX_FORCE_MU])
def test_neumann_adv_parity_fp64(shape, force):
    say("neumann advected fp64", shape, force)
    cfg = cfg_na(shape, force_form=force, eta=1.3, rho=0.8)
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


def test_neumann_adv_changes_the_step():
    """The advected chain is really taken: it differs from the plain Neumann step."""
    shape = (32, 16, 48)
    psi = structured_ic(shape)
    outs = []
    for adv in (0, 1):
        with Solver(cfg_na(shape, advect=adv)) as s:
            s.set_psi(psi)
            s.step(20)
            outs.append(s.get_psi())
    assert rel_l2(outs[1], outs[0]) > 1e-6


@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_neumann_adv_workload_small(prec):
    """Advection-coupled Neumann run: eps=0.5, q0=1, gamma=1, dt=0.02, eta=1, mu grad psi force,
    psi0 ~ U[-0.05, 0.05], 8 points per wavelength; 32^3 for 200 steps."""
    n = 32
    mk = lambda p: cfg_na((n, n, n), p, eps=0.5, q0=1.0, beta=0.0, gamma=1.0, dt=0.02, nl_model=A.SMX_NL_CUBIC,
                          force_form=A.SMX_FORCE_MU)
    psi = O.ic_noise(n, n, n, 77, 0.05)
    orc = O.Oracle(mk(A.SMX_FP64))
    orc.set_psi(psi)
    orc.step(200)
    with Solver(mk(prec)) as s:
        s.set_psi(psi)
        s.step(200)
        e = rel_l2(s.get_psi(), orc.get_psi())
        say("  neumann advected workload", prec, "rel L2", e)
        assert e < (1e-10 if prec == A.SMX_FP64 else 1e-4)


def test_neumann_adv_checkpoint_resume_exact(tmp_path):
    shape = (32, 32, 48)
    cfg = cfg_na(shape, precision=A.SMX_FP32_MIXED)
    psi = structured_ic(shape)
    path = str(tmp_path / "na.ckpt")
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(7)
        s.save_checkpoint(path)
        s.step(11)
        ref = s.get_psi()
    with Solver(cfg) as r:
        r.load_checkpoint(path)
        r.step(11)
        assert np.array_equal(r.get_psi(), ref)


def test_neumann_adv_profile_slots():
    shape = (32, 32, 32)
    with Solver(cfg_na(shape)) as s:
        s.set_psi(structured_ic(shape))
        prof = s.profile_step(3)
    names = [p[0] for p in prof]
    assert len(names) == 15 and names[4] == "k_dctz_force" and names[8] == "k_dctx_stokes" and names[-1] == "k_dctx_adv"
    assert all(p[1] > 0 for p in prof)


@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_neumann_adv_dtpsi_and_surftrack(prec):
    """dpsi/dt with the Neumann advective term, and the DCT surfTrack on the advected state."""
    shape = (32, 16, 48)
    say("neumann advected dtpsi", prec)
    cfg = cfg_na(shape, prec, force_form=A.SMX_FORCE_MU)
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
        (gz, gv, gh, gk), gf = s.surftrack(0.7)
    (oz, ov, oh, ok), of = orc.surftrack(0.7)
    say("  found", gf, of)
    assert e < 1e-10
    assert gf == of and gf > 0 and np.array_equal(gz, oz)
    np.testing.assert_allclose(gv, ov, rtol=0, atol=1e-9 * np.abs(ov).max())
