"""Uneven batch of independent boxes advanced through one captured hipGraph
(smx_batch_run) against the same boxes advanced one by one."""
import math
import sys

import numpy as np
import pytest

from oracle import oracle as O
from smectic3d_amd import Batch, Solver
from smectic3d_amd import _abi as A

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def box_cfgs(seed=5, nbox=6, sizes=(32, 48, 64), prec=A.SMX_FP64):
    rng = np.random.default_rng(seed)
    cfgs = []
    for b in range(nbox):
        n = int(rng.choice(sizes))
        L = n / 8 * 2 * math.pi  # dx = 2 pi / 8, independent of q0
        cfgs.append(A.make_config(n, lx=L, ly=L, lz=L, eps=float(rng.uniform(0.2, 0.6)),
                                  q0=float(rng.uniform(0.8, 1.2)), beta=0.0, gamma=1.0, dt=0.05,
                                  nl_model=A.SMX_NL_CUBIC, precision=prec))
    return cfgs


def assert_rounding_level(got, want, prec):
    """Grouped and single-box kernels share their tile bodies, but the compiler fuses a few
    multiply-adds differently in each (parameters from memory instead of kernel arguments), so
    the runs agree to rounding, not bit for bit."""
    tol = 1e-13 if prec == A.SMX_FP64 else 2e-6
    e = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert e < tol, e


@pytest.mark.parametrize("prec,sizes", [(A.SMX_FP64, (32, 48, 64)), (A.SMX_FP32_MIXED, (32, 48, 64)),
                                         (A.SMX_FP32_MIXED, (64, 96, 128))])
def test_batch_graph_equals_individual_runs(prec, sizes):
    cfgs = box_cfgs(prec=prec, sizes=sizes)
    psis = [O.ic_noise(c.nx, c.ny, c.nz, 100 + i, 0.05) for i, c in enumerate(cfgs)]
    say("batch sizes", [c.nx for c in cfgs])
    with Batch(cfgs) as b:
        for s, p in zip(b.boxes, psis):
            s.set_psi(p)
        b.run(23, graph_steps=5)  # 4 graph launches + 3 direct steps
        # boxes of one size share the grouped kernels: four launches per step per size
        assert b.graph_nodes() >= len({c.nx for c in cfgs}) * 5 * 4
        got = [s.get_psi() for s in b.boxes]
    for c, p, g in zip(cfgs, psis, got):
        with Solver(c) as s:
            s.set_psi(p)
            s.step(23)
            assert_rounding_level(g, s.get_psi(), prec)


@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_grouped_boxes_equal_individual_runs(prec):
    """Boxes of one grid and schedule, each with its own eps and q0, advance with one grouped
    launch per pass (4 kernel nodes per step in all): the single-box runs to rounding, and
    bit for bit the same on a second batch."""
    cfgs = box_cfgs(seed=9, nbox=7, sizes=(48,), prec=prec)
    psis = [O.ic_noise(c.nx, c.ny, c.nz, 200 + i, 0.05) for i, c in enumerate(cfgs)]
    with Batch(cfgs) as b:
        for s, p in zip(b.boxes, psis):
            s.set_psi(p)
        b.run(12, graph_steps=4)
        nodes = b.graph_nodes()
        say("  grouped graph nodes", nodes)
        assert 4 * 4 <= nodes < len(cfgs) * 4 * 4
        got = [s.get_psi() for s in b.boxes]
    with Batch(cfgs) as b:
        for s, p in zip(b.boxes, psis):
            s.set_psi(p)
        b.run(12, graph_steps=4)
        for s, g in zip(b.boxes, got):
            np.testing.assert_array_equal(s.get_psi(), g)
    for c, p, g in zip(cfgs, psis, got):
        with Solver(c) as s:
            s.set_psi(p)
            s.step(12)
            assert_rounding_level(g, s.get_psi(), prec)


def test_batch_matches_oracle_one_box():
    cfgs = box_cfgs(nbox=3)
    psis = [O.ic_noise(c.nx, c.ny, c.nz, 7 + i, 0.05) for i, c in enumerate(cfgs)]
    with Batch(cfgs) as b:
        for s, p in zip(b.boxes, psis):
            s.set_psi(p)
        b.run(40, graph_steps=10)
        got = b.boxes[1].get_psi()
    o = O.Oracle(cfgs[1])
    o.set_psi(psis[1])
    o.step(40)
    ref = o.get_psi()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-10


def test_growth_sweep_batch_matches_oracle():
    """pf3dVal harness as one GPU batch (one box per wavenumber) against the oracle's sweep
    and the growth law sigma/(1 - dt sigma/2) at lattice wavenumbers."""
    import math
    from smectic3d_amd.sweeps import growth_sweep
    cfg = A.make_config(16, 16, 32, nw=16, eps=-0.7, beta=2.0, gamma=1.0, dt=1e-3)
    dmq = 2 * math.pi / cfg.lz
    g = growth_sweep(cfg, dmq, dmq, 10)
    g0 = O.growth_sweep(cfg, dmq, dmq, 10)
    np.testing.assert_allclose(g, g0, rtol=1e-9)
    mq = dmq * np.arange(1, 11)
    sigma = -0.7 - (mq ** 2 - 1.0) ** 2
    np.testing.assert_allclose(g, sigma / (1 - 1e-3 * sigma / 2), rtol=1e-6)


@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_batch_of_every_step_kind(prec):
    """Periodic and Neumann boxes, plain and advected, in one captured graph: each equals the
    same box advanced alone, bit for bit."""
    cfgs = []
    for bc in (A.SMX_BC_PERIODIC, A.SMX_BC_NEUMANN):
        for adv in (0, 1):
            n = 32 if adv else 48
            L = n / 8 * 2 * math.pi
            cfgs.append(A.make_config(n, n, n + 16, lx=L, ly=L, lz=L * (n + 16) / n, bc=bc, eps=0.5, q0=1.0,
                                      beta=0.0, gamma=1.0, dt=0.02, nl_model=A.SMX_NL_CUBIC, advect=adv,
                                      force_form=A.SMX_FORCE_MU, precision=prec))
    psis = [O.ic_noise(c.nx, c.ny, c.nz, 300 + i, 0.05) for i, c in enumerate(cfgs)]
    say("batch of kinds", prec)
    with Batch(cfgs) as b:
        for s, p in zip(b.boxes, psis):
            s.set_psi(p)
        b.autotune(2)  # tunes every box except the advected Neumann one, which keeps its tiles
        scheds = [s.schedule() if not (c.bc == A.SMX_BC_NEUMANN and c.advect) else None
                  for s, c in zip(b.boxes, cfgs)]
        b.run(12, graph_steps=4)
        got = [s.get_psi() for s in b.boxes]
    for c, p, g, sc in zip(cfgs, psis, got, scheds):
        with Solver(c) as s:
            if sc is not None:
                s.set_schedule(sc)
            s.set_psi(p)
            s.step(12)
            np.testing.assert_array_equal(g, s.get_psi())


def test_batch_recaptures_after_a_schedule_change():
    """A box's schedule set after the capture takes it out of its group: the next run captures
    again (the stale graph would run the old tiles and the old partial counts)."""
    cfgs = box_cfgs(seed=11, nbox=3, sizes=(48,))
    psis = [O.ic_noise(c.nx, c.ny, c.nz, 500 + i, 0.05) for i, c in enumerate(cfgs)]
    alt = [(8, 0), (16, 0), (8, 0), (16, 2)]
    with Batch(cfgs) as b:
        for s, p in zip(b.boxes, psis):
            s.set_psi(p)
        b.run(4, graph_steps=2)
        n_grouped = b.graph_nodes()
        b.boxes[2].set_schedule(alt)
        b.run(4, graph_steps=2)
        assert b.graph_nodes() > n_grouped  # box 2 now runs its own branch
        assert b.boxes[2].schedule() == alt
        got = [s.get_psi() for s in b.boxes]
    for i, (c, p, g) in enumerate(zip(cfgs, psis, got)):
        with Solver(c) as s:
            s.set_psi(p)
            s.step(4)
            if i == 2:
                s.set_schedule(alt)
            s.step(4)
            assert_rounding_level(g, s.get_psi(), A.SMX_FP64)
