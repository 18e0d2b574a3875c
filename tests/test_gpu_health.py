"""Run health and state hand-over on the GPU: the non-finite checks of smx_step, L1
tracking, loading a checkpoint into a handle with queued steps, and the adjoint's memory
lent for one call only."""
import sys

import numpy as np
import pytest

from helpers import cfg_hyperboloid_256, cfg_reference, rel_l2
from oracle import oracle as O
from smectic3d_amd import Solver, SmxError
from smectic3d_amd import _abi as A

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [
    ("periodic fp64", dict(precision=A.SMX_FP64)),
    ("periodic fp32", dict(precision=A.SMX_FP32_MIXED)),
    ("neumann fp32", dict(precision=A.SMX_FP32_MIXED, bc=A.SMX_BC_NEUMANN)),
    ("advected fp64", dict(precision=A.SMX_FP64, advect=1, dt=2e-3)),
]


@pytest.mark.parametrize("name,kw", CASES)
def test_nan_state_is_reported(name, kw):
    """A NaN in psi0 spreads through the transforms (NaN arithmetic does not fault); the step
    check at step 100 flags it: smx_sync and the next smx_step return SMX_ERR_NONFINITE."""
    say("nonfinite", name)
    shape = (32, 32, 48)
    cfg = cfg_reference(*shape, **kw)
    psi = O.ic_noise(*shape, 8, 0.05)
    psi[3, 5, 7] = np.nan
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(100)  # asynchronous: queued, not yet checked on the host
        with pytest.raises(SmxError) as e:
            s.sync()
        assert e.value.code == A.SMX_ERR_NONFINITE
        with pytest.raises(SmxError) as e:
            s.step(1)
        assert e.value.code == A.SMX_ERR_NONFINITE
        with pytest.raises(SmxError) as e:
            s.l1()
        assert e.value.code == A.SMX_ERR_NONFINITE
        # a new state clears the flag
        s.set_psi(O.ic_noise(*shape, 9, 0.05))
        s.step(200)
        s.sync()


def test_nan_check_off_and_cadence():
    shape = (32, 32, 32)
    cfg = cfg_reference(*shape)
    psi = O.ic_noise(*shape, 2, 0.05)
    psi[0, 0, 0] = np.inf
    with Solver(cfg) as s:
        s.set_finite_check(0)
        s.set_psi(psi)
        s.step(150)
        s.sync()  # no check ran
        s.set_finite_check(7)
        s.set_psi(psi)
        s.step(6)
        s.sync()  # no step index is a multiple of 7 yet
        s.step(1)
        with pytest.raises(SmxError) as e:
            s.sync()
        assert e.value.code == A.SMX_ERR_NONFINITE


def test_finite_run_is_clean():
    cfg = cfg_reference(32, 32, 48, precision=A.SMX_FP32_MIXED)
    with Solver(cfg) as s:
        s.set_psi(O.ic_noise(32, 32, 48, 4, 0.05))
        for _ in range(5):
            s.step(100)
        s.sync()
        assert np.isfinite(s.l1())


def test_l1_tracking_switch_and_oracle():
    shape = (32, 32, 48)
    cfg = cfg_reference(*shape, dt=5e-3)
    psi = O.ic_noise(*shape, 6, 0.2)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.set_l1_tracking(False)
        s.step(5)
        orc.step(5)
        with pytest.raises(SmxError) as e:
            s.l1()
        assert e.value.code == A.SMX_ERR_STATE
        s.set_l1_tracking(True)
        for n in (1, 4, 13):
            s.step(n)
            orc.step(n)
            got, want = s.l1(), orc.l1()
            say("  L1 after", n, got, want)
            assert abs(got - want) < 1e-10 * abs(want)


def test_load_checkpoint_into_busy_handle(tmp_path):
    """Steps still queued on the handle's stream do not overwrite a checkpoint loaded after
    them: the resumed run equals the uninterrupted one bit for bit."""
    shape = (64, 64, 64)
    cfg = cfg_reference(*shape, precision=A.SMX_FP32_MIXED)
    psi = O.ic_noise(*shape, 12, 0.05)
    path = tmp_path / "busy.ckpt"
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(7)
        s.save_checkpoint(path)
        s.step(9)
        want = s.get_psi()
        s.step(400)  # queued, not waited for
        s.load_checkpoint(path)
        s.step(9)
        got = s.get_psi()
        assert s.step_count() == 16
    assert np.array_equal(got, want)


def test_adjoint_memory_is_lent_for_one_call():
    """The default adjoint budget may take most of the HBM for the sweep (256^3 x 2000 steps
    needs more than the device holds), but returns it: surfTrack's scratch on the same handle,
    a second 256^3 handle and its steps allocate afterwards."""
    n, T = 256, 2000
    cfg = cfg_hyperboloid_256(precision=A.SMX_FP32_MIXED)
    cfg.eps, cfg.dt = 0.3, 0.05
    psi = O.ic_noise(n, n, n, 3, 0.05)
    with Solver(cfg) as s:
        J, d, de, dq = s.adjoint_grad(psi, T, 50)
        res, rec, nbytes = s.adjoint_info()
        say("  adjoint resident", res, "recomputed", rec, "GB", nbytes / 1e9)
        assert nbytes > 100e9 and np.isfinite(de) and np.isfinite(dq)
        (sz, v, H, K), found = s.surftrack(0.02)  # arms inside the noise of amplitude 0.05
        assert found > 0
        with Solver(cfg) as t:
            t.set_psi(psi)
            t.step(10)
            t.sync()
            s.step(10)
            assert rel_l2(t.get_psi(), s.get_psi()) == 0.0
