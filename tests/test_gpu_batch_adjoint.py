"""Batched gradients (smx_batch_adjoint_grad): the adjoint sweeps of uneven boxes as
parallel branches of one captured hipGraph, against the single-box smx_adjoint_grad and against
the FP64 oracle's adjoint.  A box alone in its branch runs the single-box kernels (bit for bit
the same gradient); boxes of one grid sweep together through the grouped kernels, whose
tile bodies are the same but whose multiply-adds the compiler fuses a little differently, so
their gradients agree to rounding."""
import math
import sys

import numpy as np
import pytest

from helpers import rel_l2
from oracle import oracle as O
from smectic3d_amd import Batch, Solver
from smectic3d_amd import _abi as A

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def box_cfgs(sizes, seed=5, prec=A.SMX_FP64):
    """The uneven-batch workload: eps ~ U[0.2, 0.6], q0 ~ U[0.8, 1.2], gamma = 1, dt = 0.05,
    dx = 2 pi / 8 independent of q0 (so that d/dq0 does not rescale the grid)."""
    rng = np.random.default_rng(seed)
    cfgs = []
    for n in sizes:
        L = n / 8 * 2 * math.pi
        cfgs.append(A.make_config(n, lx=L, ly=L, lz=L, eps=float(rng.uniform(0.2, 0.6)),
                                  q0=float(rng.uniform(0.8, 1.2)), beta=0.0, gamma=1.0, dt=0.05,
                                  nl_model=A.SMX_NL_CUBIC, precision=prec))
    return cfgs


def single(cfg, psi, T, K, amp):
    with Solver(cfg) as s:
        return s.adjoint_grad(psi, T, K, amp)


@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
def test_batch_adjoint_equals_single_box(prec):
    sizes = [32, 64, 48, 96, 128, 64]
    cfgs = box_cfgs(sizes, prec=prec)
    T, K = 30, 10
    psis = [O.ic_noise(c.nx, c.ny, c.nz, 40 + i, 0.05) for i, c in enumerate(cfgs)]
    amps = [0.5, 0.4, 0.5, 0.3, 0.5, 0.45]
    say("batch adjoint", sizes, "prec", prec)
    with Batch(cfgs) as b:
        J, d, de, dq = b.adjoint_grad(psis, T, K, amps)
        nodes = b.adjoint_nodes()
        after = [s.get_psi() for s in b.boxes]
    say("  graph nodes", nodes)
    assert nodes > len(set(sizes)) * 8 * T  # one branch per grid
    tol = 1e-12 if prec == A.SMX_FP64 else 1e-5
    for i, (c, p) in enumerate(zip(cfgs, psis)):
        J1, d1, e1, q1 = single(c, p, T, K, amps[i])
        say("  box", i, c.nx, "J", J[i], J1, "deps", de[i], e1, "dq0", dq[i], q1)
        if sizes.count(c.nx) == 1:
            assert J[i] == J1 and de[i] == e1 and dq[i] == q1
            assert np.array_equal(d[i], d1)
        else:
            assert abs(J[i] - J1) <= tol * abs(J1)
            assert abs(de[i] - e1) <= tol * abs(e1) and abs(dq[i] - q1) <= tol * abs(q1)
            assert rel_l2(d[i], d1) < tol
        assert rel_l2(after[i], p) < (1e-13 if prec == A.SMX_FP64 else 1e-6)  # every box is left at psi0


def test_batch_adjoint_matches_oracle():
    sizes = [32, 48, 96]
    cfgs = box_cfgs(sizes, seed=9)
    T, K = 12, 5
    psis = [O.ic_noise(c.nx, c.ny, c.nz, 70 + i, 0.05) for i, c in enumerate(cfgs)]
    with Batch(cfgs) as b:
        J, d, de, dq = b.adjoint_grad(psis, T, K, 0.5)
    for i, (c, p) in enumerate(zip(cfgs, psis)):
        J0, d0, e0, q0 = O.adjoint_grad(c, p, T, 0.5)
        say("  box", c.nx, "deps", de[i], e0, "dq0", dq[i], q0)
        assert abs(J[i] - J0) < 1e-12 * abs(J0)
        assert abs(de[i] - e0) < 1e-9 * max(1, abs(e0)) and abs(dq[i] - q0) < 1e-9 * max(1, abs(q0))
        assert np.linalg.norm(d[i] - d0) < 1e-9 * np.linalg.norm(d0)


def test_batch_adjoint_budgets_and_graph_reuse():
    """Explicit budgets: one record set per box (every segment but one recomputed) gives the
    same gradients bit for bit; with every box keeping its workspace the captured graph is
    reused by the next call and recaptured when a target amplitude changes."""
    sizes = [64, 96, 64, 128]
    cfgs = box_cfgs(sizes, seed=3, prec=A.SMX_FP32_MIXED)
    T, K = 40, 10
    psis = [O.ic_noise(c.nx, c.ny, c.nz, 90 + i, 0.05) for i, c in enumerate(cfgs)]
    with Batch(cfgs) as b:
        ref = b.adjoint_grad(psis, T, K, 0.5)  # default budgets: everything resident
        for s, c in zip(b.boxes, cfgs):
            n = c.nx * c.ny * c.nz
            setbytes = 4 * n * K + 8 * c.nx * c.ny * (c.nz // 2 + 1) * (K + 1)
            s.set_adjoint_budget(setbytes)  # one resident segment of four
        got = b.adjoint_grad(psis, T, K, 0.5)
        for s in b.boxes:
            res, rec, _ = s.adjoint_info()
            assert (res, rec) == (K, T - K)
        nodes1 = b.adjoint_nodes()
        again = b.adjoint_grad(psis, T, K, 0.5)
        assert b.adjoint_nodes() == nodes1
        moved = b.adjoint_grad(psis, T, K, [0.5, 0.5, 0.5, 0.2])
    for i in range(len(sizes)):
        assert got[0][i] == ref[0][i] and got[2][i] == ref[2][i] and got[3][i] == ref[3][i]
        assert np.array_equal(got[1][i], ref[1][i]) and np.array_equal(again[1][i], ref[1][i])
        assert again[2][i] == ref[2][i] and again[3][i] == ref[3][i]
    assert moved[2][:3] == ref[2][:3] and moved[2][3] != ref[2][3]
