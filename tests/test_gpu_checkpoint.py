"""Exact checkpoints (smx_save_checkpoint / smx_load_checkpoint): a run resumed from a
checkpoint continues bit for bit, and foreign or damaged files are refused."""
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


CASES = [
    ("periodic fp64", dict(precision=A.SMX_FP64)),
    ("periodic fp32", dict(precision=A.SMX_FP32_MIXED)),
    ("neumann fp32", dict(precision=A.SMX_FP32_MIXED, bc=A.SMX_BC_NEUMANN)),
    ("advected fp64", dict(precision=A.SMX_FP64, advect=1, dt=2e-3)),
]


@pytest.mark.parametrize("name,kw", CASES)
def test_resume_is_bit_exact(tmp_path, name, kw):
    say("checkpoint", name)
    shape = (32, 32, 48)
    cfg = cfg_reference(*shape, **kw)
    psi = O.ic_noise(*shape, 8, 0.05)
    path = tmp_path / "state.ckpt"
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(7)
        s.save_checkpoint(path)
        s.step(9)
        want = s.get_psi()
        e_want = s.energy()
        assert np.isfinite(want).all() and np.isfinite(e_want)
        assert s.step_count() == 16
    with Solver(cfg) as r:
        r.load_checkpoint(path)
        assert r.step_count() == 7
        r.step(9)
        assert r.step_count() == 16
        np.testing.assert_array_equal(r.get_psi(), want)
        assert r.energy() == e_want


def test_foreign_and_damaged_checkpoints_refused(tmp_path):
    shape = (32, 32, 48)
    cfg = cfg_reference(*shape)
    path = tmp_path / "state.ckpt"
    with Solver(cfg) as s:
        s.set_psi(O.ic_noise(*shape, 8, 0.3))
        s.step(3)
        s.save_checkpoint(path)
    other = cfg_reference(*shape, eps=0.1)
    with Solver(other) as s2:
        with pytest.raises(SmxError) as ei:
            s2.load_checkpoint(path)
        assert ei.value.code == A.SMX_ERR_INVALID
    raw = bytearray(path.read_bytes())
    raw[len(raw) // 2] ^= 0x40
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(bytes(raw))
    with Solver(cfg) as s3:
        with pytest.raises(SmxError) as ei:
            s3.load_checkpoint(bad)
        assert ei.value.code == A.SMX_ERR_INVALID
        with pytest.raises(SmxError):
            s3.load_checkpoint(tmp_path / "missing.ckpt")
