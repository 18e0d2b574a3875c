"""Tile schedules and the autotuner of the periodic and Neumann steps (libsmectic through the
C ABI).

Every schedule candidate (periodic: columns per tile 8/16/32, radix-plan variant 0/1, and for
the x pass the carry-order forms 2-7; Neumann: tile widths 32/16 in FP32, 16/8 in FP64) of
every step kernel must give the oracle's result to the usual tolerance: FP64 relative
L2(psi) <= 1e-11 after 5 steps, FP32-mixed <= 2e-5.  The autotuner must leave the
state untouched, return a valid schedule, and the tuned step must keep parity.
"""
import sys

import numpy as np
import pytest

from helpers import cfg_plain_64, cfg_reference, rel_l2
from oracle import oracle as O
from smectic3d_amd import Batch, Solver, SmxError
from smectic3d_amd import _abi as A

pytestmark = pytest.mark.gpu

NCS = (8, 16, 32)
VARIANTS = (0, 1)
VARIANTS_X = (0, 1, 2, 3, 4, 5, 6, 7)
DEFAULT = [(16, 0)] * 4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# every power-of-two plan length appears in each axis role
SHAPES = [(64, 64, 64), (128, 32, 256), (256, 16, 128), (512, 16, 32), (32, 512, 16), (16, 32, 512)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_every_schedule_candidate_matches_oracle(shape, prec):
    say("schedules", shape, prec)
    cfg = cfg_reference(*shape, eps=0.5, beta=0.3, dt=0.05, nl_model=A.SMX_NL_CUBIC, precision=prec)
    psi = O.ic_noise(*shape, 23, 0.3)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    orc.step(5)
    want = orc.get_psi()
    tol = 1e-11 if prec == A.SMX_FP64 else 2e-5
    ran = 0
    with Solver(cfg) as s:
        assert s.schedule() == DEFAULT
        for slot in range(4):
            for nc in NCS:
                for v in (VARIANTS_X if slot == 3 else VARIANTS):
                    sched = list(DEFAULT)
                    sched[slot] = (nc, v)
                    try:
                        s.set_schedule(sched)
                    except SmxError as e:
                        assert e.code == A.SMX_ERR_UNSUPPORTED
                        continue
                    assert s.schedule() == sched
                    s.set_psi(psi)
                    s.step(5)
                    err = rel_l2(s.get_psi(), want)
                    say("  slot", slot, "nc", nc, "variant", v, "err", err)
                    assert err < tol, (slot, nc, v, err)
                    ran += 1
        # all slots off the default at once
        alt = [(8, 1), (32, 1), (8, 1), (8, 1)]
        try:
            s.set_schedule(alt)
        except SmxError:
            alt = None
        if alt is not None:
            s.set_psi(psi)
            s.step(5)
            assert rel_l2(s.get_psi(), want) < tol
    assert ran >= 12  # the default plus at least two alternatives per slot


NEU_SHAPES = [(64, 64, 64), (32, 48, 96), (96, 32, 48), (48, 96, 32), (16, 32, 384)]


@pytest.mark.parametrize("shape", NEU_SHAPES)
@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_every_neumann_tile_width_matches_oracle(shape, prec):
    say("neumann widths", shape, prec)
    cfg = cfg_reference(*shape, bc=A.SMX_BC_NEUMANN, eps=0.5, beta=0.3, dt=0.05, nl_model=A.SMX_NL_CUBIC,
                        precision=prec)
    psi = O.ic_noise(*shape, 29, 0.3)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    orc.step(5)
    want = orc.get_psi()
    tol = 1e-11 if prec == A.SMX_FP64 else 2e-5
    wide, narrow = (32, 16) if prec == A.SMX_FP32_MIXED else (16, 8)
    ran = 0
    with Solver(cfg) as s:
        base = s.schedule()
        assert base == [(wide, 0)] * 4 or base[3] == (16, 0)  # x falls back to 16 when ny is not a multiple
        for sched in ([(narrow, 0)] + base[1:], base[:1] + [(narrow, 0)] + base[2:],
                      base[:2] + [(narrow, 0)] + base[3:], base[:3] + [(narrow, 0)], [(narrow, 0)] * 4):
            try:
                s.set_schedule(sched)
            except SmxError as e:
                assert e.code == A.SMX_ERR_UNSUPPORTED
                continue
            s.set_psi(psi)
            s.step(5)
            err = rel_l2(s.get_psi(), want)
            say("  ", sched, "err", err)
            assert err < tol, (sched, err)
            ran += 1
    assert ran >= 4


@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_neumann_autotune_keeps_state_and_parity(prec):
    shape = (64, 48, 96)
    cfg = cfg_reference(*shape, bc=A.SMX_BC_NEUMANN, eps=0.5, beta=0.3, dt=0.05, nl_model=A.SMX_NL_CUBIC,
                        precision=prec)
    psi = O.ic_noise(*shape, 37, 0.05)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    orc.step(20)
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(10)
        before = s.get_psi()
        chosen = s.autotune(5)
        say("  neumann chosen", chosen)
        assert s.schedule() == chosen and all(v == 0 for _, v in chosen)
        assert np.array_equal(s.get_psi(), before)
        s.step(10)
        err = rel_l2(s.get_psi(), orc.get_psi())
    assert err < (1e-11 if prec == A.SMX_FP64 else 1e-5)


def test_invalid_schedules_rejected():
    say("invalid schedules")
    with Solver(cfg_reference(64, 48, 32)) as s:
        for bad in ([(7, 0)] + DEFAULT[1:], [(16, 2)] + DEFAULT[1:], DEFAULT[:3] + [(32, 0)],
                    DEFAULT[:3] + [(16, 8)]):
            with pytest.raises(SmxError) as ei:
                s.set_schedule(bad)  # nc 7; y variant 2; x tiles of 32 do not divide ny = 48; x variant 8
            assert ei.value.code == A.SMX_ERR_UNSUPPORTED
        assert s.schedule() == DEFAULT
    with Solver(cfg_reference(32, 32, 32, bc=A.SMX_BC_NEUMANN)) as s:  # FP64 Neumann: widths 16 or 8
        assert s.schedule() == [(16, 0)] * 4
        for bad in ([(32, 0)] + [(16, 0)] * 3, [(16, 1)] + [(16, 0)] * 3):
            with pytest.raises(SmxError) as ei:
                s.set_schedule(bad)  # no 32-wide FP64 Neumann y pass; no second radix plan
            assert ei.value.code == A.SMX_ERR_UNSUPPORTED
    with Solver(cfg_reference(32, 32, 32, bc=A.SMX_BC_NEUMANN, advect=1, dt=2e-3)) as s:
        with pytest.raises(SmxError) as ei:
            s.set_schedule([(8, 0)] + [(16, 0)] * 3)  # the advected Neumann step keeps its tiles
        assert ei.value.code == A.SMX_ERR_UNSUPPORTED
        assert s.autotune(2) == [(16, 0)] * 4


@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_autotune_keeps_state_and_parity(prec):
    say("autotune 64^3", prec)
    cfg = cfg_plain_64(precision=prec)
    shape = (64, 64, 64)
    psi = O.ic_noise(*shape, 31, 0.01)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    orc.step(20)
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(10)
        before = s.get_psi()
        e_before = s.energy()
        chosen = s.autotune(5)
        say("  chosen", chosen)
        assert s.schedule() == chosen
        assert all(nc in NCS and v in (VARIANTS_X if k == 3 else VARIANTS) for k, (nc, v) in enumerate(chosen))
        assert np.array_equal(s.get_psi(), before)  # tuning ran on scratch buffers only
        assert s.energy() == e_before
        s.step(10)
        err = rel_l2(s.get_psi(), orc.get_psi())
    tol = 1e-11 if prec == A.SMX_FP64 else 1e-5
    assert err < tol
    # a second handle of the same shape reuses the cached tuning
    with Solver(cfg) as s2:
        assert s2.autotune(5) == chosen


def test_batch_autotune_recaptures_graph():
    say("batch autotune")
    shapes = [64, 96, 64, 128]
    cfgs = [cfg_plain_64(precision=A.SMX_FP32_MIXED, n=n) for n in shapes]
    psis = [O.ic_noise(n, n, n, 40 + i, 0.05) for i, n in enumerate(shapes)]
    with Batch(cfgs) as b, Batch(cfgs) as ref:
        for box, r, p in zip(b.boxes, ref.boxes, psis):
            box.set_psi(p)
            r.set_psi(p)
        b.run(10, graph_steps=5)
        scheds = b.autotune(3)
        say("  schedules", scheds)
        assert len(scheds) == len(shapes)
        assert scheds[0] == scheds[2]  # same shape, same (cached) tuning
        b.run(10, graph_steps=5)  # captured again with the tuned schedules
        assert b.graph_nodes() >= 4 * 5 * len(set(shapes))  # one grouped launch per pass and shape
        ref.run(20, graph_steps=5)
        for box, r in zip(b.boxes, ref.boxes):
            assert rel_l2(box.get_psi(), r.get_psi()) < 1e-5
