"""smectic3d_amd -- MI355X (gfx950) implementation of the semi-implicit
pseudo-spectral Swift-Hohenberg time step of evitral/smectic3D.

The compute path is libsmectic.so (hand-written HIP kernels behind the C ABI
in include/smectic.h); this package is the host-side mirror of the reference
loop over that ABI.
"""
from ._abi import *  # noqa: F401,F403
from ._lib import EXPORTS, SmxError, lib  # noqa: F401
from .solver import Batch, Solver  # noqa: F401
