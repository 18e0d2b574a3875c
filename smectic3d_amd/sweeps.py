"""Batched physics sweeps on the hipGraph batch path.

`growth_sweep` is the pf3dVal harness (periodicFFT/pf3dVal.cpp:214-329) run as one batch:
the reference runs one MPI job per wavenumber; here every wavenumber is a box of the same
grid, all boxes advance together through one captured hipGraph, and the growth rate is
read at the box centre exactly as the reference reads it (:316-320).
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
