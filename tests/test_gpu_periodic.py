"""GPU parity of the periodic Swift-Hohenberg step (libsmectic through the C ABI)
against the FP64 CPU oracle on identical seeded inputs.

Tolerances (FP64 path): relative L2(psi) <= 1e-10, relative energy <= 1e-12.
FP32-mixed path: relative L2(psi) <= 1e-5 over the short runs checked here
(the error of the mixed path against the FP64 oracle is also reported by the
benchmark at the full 256^3 size).
"""
import math
import sys

import numpy as np
import pytest

from helpers import cfg_plain_64, cfg_reference, rel_l2
from oracle import oracle as O
from smectic3d_amd import Solver, SmxError
from smectic3d_amd import _abi as A

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noise(shape, seed, amp):
    return O.ic_noise(*shape, seed, amp)


SHAPES = [(16, 16, 16), (32, 48, 24), (24, 32, 64), (48, 16, 96), (64, 64, 64)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_set_get_roundtrip(shape, prec):
    say("roundtrip", shape, prec)
    cfg = cfg_reference(*shape, precision=prec)
    psi = noise(shape, 3, 0.5)
    with Solver(cfg) as s:
        s.set_psi(psi)
        out = s.get_psi()
    tol = 1e-14 if prec == A.SMX_FP64 else 2e-6
    assert rel_l2(out, psi) < tol


def test_padded_row_stride():
    say("row_stride")
    shape = (16, 32, 24)
    cfg = cfg_reference(*shape)
    psi = noise(shape, 5, 0.3)
    rs = 2 * (shape[2] // 2 + 1)  # FFTW padded r2c layout, pf3d.cpp:312
    flat = np.full(shape[0] * shape[1] * rs, 7.0)
    flat.reshape(shape[0] * shape[1], rs)[:, : shape[2]] = psi.reshape(-1, shape[2])
    with Solver(cfg) as s:
        s.set_psi(flat, row_stride=rs)
        out = s.get_psi(row_stride=rs).reshape(shape[0] * shape[1], rs)
    assert rel_l2(out[:, : shape[2]], psi.reshape(-1, shape[2])) < 1e-14
    assert np.all(out[:, shape[2]:] == 0.0)  # padding untouched (zero-initialised)


@pytest.mark.parametrize("shape", SHAPES[:4])
@pytest.mark.parametrize("model", [A.SMX_NL_QUINTIC, A.SMX_NL_CUBIC])
def test_step_parity_fp64(shape, model):
    say("step parity fp64", shape, model)
    if model == A.SMX_NL_QUINTIC:
        cfg = cfg_reference(*shape, dt=5e-3)
        amp = 0.8
    else:
        cfg = cfg_reference(*shape, eps=0.5, beta=0.3, dt=0.05, nl_model=A.SMX_NL_CUBIC)
        amp = 0.2
    psi = noise(shape, 11, amp)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    with Solver(cfg) as s:
        s.set_psi(psi)
        for n in (1, 7, 30):
            s.step(n)
            orc.step(n)
            e = rel_l2(s.get_psi(), orc.get_psi())
            say("  after", s.step_count(), "steps rel L2", e)
            assert e < 1e-10
        assert abs(s.l1() - orc.l1()) / orc.l1() < 1e-9
        ge, oe = s.energy(), orc.energy()
        assert abs(ge - oe) / abs(oe) < 1e-12
        assert rel_l2(s.dtpsi(), orc.dtpsi()) < 1e-10
        gq, oq = s.dominant_wavenumber(), orc.dominant_wavenumber()
        assert abs(gq - oq) / oq < 1e-12


def test_mid_plane_penalty_fp64():
    """pf3dMorph.cpp:456-462,572-581: 2 wMid d_zz psi added to N at k = nz/2."""
    shape = (16, 32, 32)
    say("penalty", shape)
    cfg = cfg_reference(*shape, dt=2e-3, w_mid=1.0, mid_plane=shape[2] // 2)
    psi = noise(shape, 21, 0.6)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(20)
        orc.step(20)
        assert rel_l2(s.get_psi(), orc.get_psi()) < 1e-10


def test_plain_64_parity_fp64():
    """64^3 plain periodic SH (eps=0.5, beta=0, gamma=1, dt=0.05, cubic), 1000 steps."""
    say("plain 64^3 x 1000 fp64")
    cfg = cfg_plain_64()
    psi = noise((64, 64, 64), 2024, 0.01)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(1000)
        orc.step(1000)
        e = rel_l2(s.get_psi(), orc.get_psi())
        ge, oe = s.energy(), orc.energy()
        say("  rel L2", e, "rel energy", abs(ge - oe) / abs(oe))
        assert e < 1e-10
        assert abs(ge - oe) / abs(oe) < 1e-12


def test_plain_64_fp32_mixed_error():
    say("plain 64^3 x 200 fp32-mixed")
    cfg64 = cfg_plain_64()
    cfg32 = cfg_plain_64(precision=A.SMX_FP32_MIXED)
    psi = noise((64, 64, 64), 2024, 0.01)
    orc = O.Oracle(cfg64)
    orc.set_psi(psi)
    orc.step(200)
    with Solver(cfg32) as s:
        s.set_psi(psi)
        s.step(200)
        e = rel_l2(s.get_psi(), orc.get_psi())
        ge, oe = s.energy(), orc.energy()
        say("  fp32-mixed rel L2", e, "rel energy", abs(ge - oe) / abs(oe))
        assert e < 1e-5


def test_errors():
    say("errors")
    with pytest.raises(SmxError):
        Solver(A.make_config(20))  # no radix plan for 20
    cfg = cfg_reference(16, 16, 16)
    with Solver(cfg) as s:
        with pytest.raises(SmxError):
            s.step(1)  # no state yet
        with pytest.raises(SmxError):
            s.get_psi()


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
def test_every_length_parity(shape, prec):
    """Every radix plan in every axis role (x pass, y pass, z pass) against the oracle."""
    say("length parity", shape, prec)
    cfg = cfg_reference(*shape, eps=0.5, beta=0.3, dt=0.05, nl_model=A.SMX_NL_CUBIC, precision=prec)
    psi = noise(shape, 17, 0.3)
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


def test_long_run_hermitian_planes():
    """2000 steps of the hyperboloid workload at 64^3 (dx = 2pi/8, eps=0.3, dt=0.1, cubic).

    Regression for the carried CNAB2 state: the kz = 0 and kz = nz/2 planes must
    stay Hermitian, or their invisible anti-Hermitian part grows in the unstable
    band and the energy diverges while psi still looks right."""
    import bench
    say("long run 64^3 x 2000 hyperboloid")
    n = 64
    dx = (n / 8 * 2 * math.pi) / n
    psi = bench.hyperboloid_ic(n, dx)
    orc = O.Oracle(bench.workload_cfg(A.SMX_FP64, n))
    orc.set_psi(psi)
    orc.step(2000)
    oe, op = orc.energy(), orc.get_psi()
    with Solver(bench.workload_cfg(A.SMX_FP64, n)) as s:
        s.set_psi(psi)
        s.step(2000)
        e, ge = rel_l2(s.get_psi(), op), s.energy()
        say("  fp64 rel L2", e, "rel energy", abs(ge - oe) / abs(oe))
        assert e < 1e-10 and abs(ge - oe) / abs(oe) < 1e-12
    with Solver(bench.workload_cfg(A.SMX_FP32_MIXED, n)) as s:
        s.set_psi(psi)
        s.step(2000)
        e, ge = rel_l2(s.get_psi(), op), s.energy()
        say("  fp32-mixed rel L2", e, "rel energy", abs(ge - oe) / abs(oe))
        assert np.isfinite(ge) and abs(ge - oe) / abs(oe) < 1e-3
