"""Batched physics sweeps on the hipGraph batch path.

`growth_sweep` is the pf3dVal harness (periodicFFT/pf3dVal.cpp:214-329) run as one batch:
the reference runs one MPI job per wavenumber; here every wavenumber is a box of the same
grid, all boxes advance together through one captured hipGraph, and the growth rate is
read at the box centre exactly as the reference reads it (:316-320).

`eckhaus_scan` is the energyPsi.m scan (:108-234): the energies of a state perturbed by
0.01 dpsiz cos((1 + Q) z) for Q = -dq, 0, dq, ... <= qmax (the first entry unperturbed),
all evaluated on the GPU-resident field, four per pass.
"""
import copy

import numpy as np

from .solver import Batch


def growth_sweep(cfg, mq0, dmq, nq, amp=1e-4):
    """Growth rates (psi_3 - psi_2) / (dt psi_2) at the box centre for psi0 = amp cos(mq k dz),
    mq = mq0 + i dmq, i < nq, after the reference's three steps (pf3dVal.cpp:214-329; the
    CNAB2 bootstrap makes step 1 a forward-Euler step, as there).  cfg: an SmxConfig
    (periodic, no advection); returns an array of nq rates."""
    if cfg.advect:
        raise ValueError("the pf3dVal harness has no advection")
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    dz = cfg.lz / nz
    z = np.arange(nz) * dz
    cfgs = [copy.copy(cfg) for _ in range(nq)]
    centre = (nx // 2, ny // 2, nz // 2)
    with Batch(cfgs) as b:
        for i, box in enumerate(b.boxes):
            psi = np.broadcast_to(amp * np.cos((mq0 + i * dmq) * z), (nx, ny, nz))
            box.set_psi(np.ascontiguousarray(psi))
        b.run(2, graph_steps=2)
        p2 = np.array([box.get_psi()[centre] for box in b.boxes])
        b.run(1, graph_steps=1)
        p3 = np.array([box.get_psi()[centre] for box in b.boxes])
    return (p3 - p2) / (cfg.dt * p2)


def eckhaus_scan_q(lz, nz, qmax=1.4):
    """The scan's Q list and amplitudes, energyPsi.m:25,108-124 (dq = 2 pi/L; Q = -dq:dq:qmax,
    the first entry unperturbed, the others at amplitude 0.01)."""
    dq = 2 * np.pi / lz
    q = -dq + dq * np.arange(int(np.floor((qmax + dq) / dq + 1e-9)) + 1)
    amp = np.full(len(q), 0.01)
    amp[0] = 0.0
    return q, amp


def eckhaus_scan(solver, qmax=1.4):
    """(Q, energies) of the energyPsi.m scan about the solver's current state."""
    q, amp = eckhaus_scan_q(solver.cfg.lz, solver.cfg.nz, qmax)
    return q, solver.eckhaus_scan(q, amp)
