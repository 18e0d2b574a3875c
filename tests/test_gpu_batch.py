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


@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_batch_graph_equals_individual_runs(prec):
    cfgs = box_cfgs(prec=prec)
    psis = [O.ic_noise(c.nx, c.ny, c.nz, 100 + i, 0.05) for i, c in enumerate(cfgs)]
    say("batch sizes", [c.nx for c in cfgs])
    with Batch(cfgs) as b:
        for s, p in zip(b.boxes, psis):
            s.set_psi(p)
        b.run(23, graph_steps=5)  # 4 graph launches + 3 direct steps
        assert b.graph_nodes() >= len(cfgs) * 5 * 4
        got = [s.get_psi() for s in b.boxes]
    for c, p, g in zip(cfgs, psis, got):
        with Solver(c) as s:
            s.set_psi(p)
            s.step(23)
            np.testing.assert_array_equal(g, s.get_psi())


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
