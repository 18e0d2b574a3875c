"""GPU surfTrack (periodic: surfTrack.cpp:767-806; Neumann: cosNoAdvMorphCurv.cpp:811-966)
against the FP64 oracle on identical states."""
import math
import sys

import numpy as np
import pytest

from helpers import rel_l2
from oracle import oracle as O
from smectic3d_amd import Solver
from smectic3d_amd import _abi as A

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _compare(s, orc, arm, vsign):
    (gz, gv, gh, gk), gf = s.surftrack(arm, vsign)
    (oz, ov, oh, ok), of = orc.surftrack(arm, vsign)
    say("  found", gf, of)
    assert gf == of and gf > 0
    np.testing.assert_array_equal(gz, oz)
    for a, b in ((gv, ov), (gh, oh), (gk, ok)):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12 * np.abs(b).max())


def test_surftrack_neumann_matches_oracle():
    nx, ny, nz = 32, 32, 48
    dx = 2 * math.pi / 8
    cfg = A.make_config(nx, ny, nz, bc=A.SMX_BC_NEUMANN, nw=8, eps=0.5, beta=0.0, gamma=1.0, dt=0.05,
                        nl_model=A.SMX_NL_CUBIC)
    psi = O.ic_hyperboloid(nx, ny, nz, dx, dx, dx, 1.0, 0.9, 1e-3, 0)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    orc.step(40)
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(40)
        say("neumann surftrack")
        _compare(s, orc, 0.7, -1.0)


def test_surftrack_periodic_matches_oracle():
    nx, ny, nz = 32, 32, 48
    dx = 2 * math.pi / 8
    cfg = A.make_config(nx, ny, nz, nw=8, eps=0.5, beta=0.0, gamma=1.0, dt=0.05, nl_model=A.SMX_NL_CUBIC)
    psi = O.ic_hyperboloid(nx, ny, nz, dx, dx, dx, 1.0, 0.9, 1e-3, 0)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    orc.step(40)
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(40)
        say("periodic surftrack")
        _compare(s, orc, 0.1, 1.0)


def test_surftrack_columns_without_crossing_keep_previous_values():
    nx, ny, nz = 16, 16, 32
    cfg = A.make_config(nx, ny, nz, bc=A.SMX_BC_NEUMANN, lx=8.0, ly=8.0, lz=16.0, eps=0.3, beta=0.0,
                        gamma=1.0, nl_model=A.SMX_NL_CUBIC)
    z = (np.arange(nz) + 0.5) * 0.5
    psi = np.broadcast_to(np.tanh((z - 6.0) / 1.5), (nx, ny, nz)).copy()
    with Solver(cfg) as s:
        s.set_psi(psi)
        (z1, v1, h1, k1), f1 = s.surftrack(0.7)
        assert f1 == nx * ny
        s.set_psi(-np.ones((nx, ny, nz)))  # no column arms
        (z2, v2, h2, k2), f2 = s.surftrack(0.7)
        assert f2 == 0
        np.testing.assert_array_equal(z2, z1)  # persistent values, as the reference's surf arrays
        # the same values written into caller-owned arrays (out=), in place
        bufs = np.full((4, nx * ny), 7.0)
        outs, f3 = s.surftrack(0.7, out=bufs)
        assert f3 == 0 and all(np.shares_memory(o, b) for o, b in zip(outs, bufs))
        for b, ref in zip(bufs, (z2, v2, h2, k2)):
            np.testing.assert_array_equal(b, ref)
        with pytest.raises(ValueError):
            s.surftrack(0.7, out=[np.zeros(nx * ny)] * 3)


@pytest.mark.parametrize("bc,model,beta,w_mid,mid", [
    (A.SMX_BC_PERIODIC, A.SMX_NL_CUBIC, 0.3, 0.0, -1),
    (A.SMX_BC_PERIODIC, A.SMX_NL_QUINTIC, 0.5, 0.3, 14),
    (A.SMX_BC_NEUMANN, A.SMX_NL_QUINTIC, 0.5, 0.0, -1),
])
def test_surftrack_rate_with_nonlinearity(bc, model, beta, w_mid, mid):
    """The velocity's dpsi/dt = F^-1[aLin psi_hat] + N(psi) with N taken at the crossing: both
    models with beta != 0, and the mid-plane penalty of N on the plane most columns cross."""
    nx, ny, nz = 32, 32, 48
    dx = 2 * math.pi / 8
    cfg = A.make_config(nx, ny, nz, bc=bc, nw=8, eps=0.5, beta=beta, gamma=1.0, dt=0.05, nl_model=model)
    cfg.w_mid, cfg.mid_plane = w_mid, mid
    psi = O.ic_hyperboloid(nx, ny, nz, dx, dx, dx, 1.0, 0.9, 1e-3, 0)
    orc = O.Oracle(cfg)
    orc.set_psi(psi)
    orc.step(40)
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(40)
        say("surftrack rate", bc, model, beta, w_mid, mid)
        arm, vsign = (0.1, 1.0) if bc == A.SMX_BC_PERIODIC else (0.7, -1.0)
        (gz, _, _, _), _ = s.surftrack(arm, vsign)
        if mid > 0:  # the evaluation index is the crossing or the one below it
            assert np.count_nonzero((gz == mid) | (gz == mid + 1)) >= 8
        _compare(s, orc, arm, vsign)


@pytest.mark.parametrize("prec", [A.SMX_FP64, A.SMX_FP32_MIXED])
@pytest.mark.parametrize("bc", [A.SMX_BC_PERIODIC, A.SMX_BC_NEUMANN])
@pytest.mark.parametrize("shape", [(32, 48, 24), (48, 32, 64), (16, 16, 96), (16, 96, 48), (24, 128, 32),
                                   (32, 16, 192), (16, 32, 128), (16, 384, 24), (384, 16, 16), (512, 16, 24),
                                   (16, 512, 16), (16, 16, 512), (16, 16, 384)])
def test_surftrack_shapes_and_precisions(shape, bc, prec):
    """Every axis length role and both working precisions.  The field values come from one
    kernel per field that inverts y and sums the z transform at each column's index over kz
    splits (up to 16 of them for a short x axis): here the splits and the rounds of pencils
    within them come out uneven and the last round is partly idle.  A grid whose z blocks are whole
    waves (Neumann nz >= 96, periodic nz >= 64) scans its columns and sums psi_zz or psi_z inside
    the z inverse; a shorter one runs the separate scan and point kernels: both forms are here, and every kernel at the longest axis lengths (384 and
    512 along each axis).  The oracle evaluates the GPU's own current field
    (get_psi), so an FP32-mixed state is compared on identical data."""
    nx, ny, nz = shape
    dx = 2 * math.pi / 8
    cfg = A.make_config(nx, ny, nz, bc=bc, nw=8, eps=0.5, beta=0.3, gamma=1.0, dt=0.05,
                        nl_model=A.SMX_NL_CUBIC, precision=prec)
    psi = O.ic_hyperboloid(nx, ny, nz, dx, dx, dx, 1.0, 0.9, 1e-3, 0)
    arm = 0.1 if bc == A.SMX_BC_PERIODIC else 0.7
    say("surftrack", shape, bc, prec)
    with Solver(cfg) as s:
        s.set_psi(psi)
        s.step(30)
        orc = O.Oracle(cfg)
        orc.set_psi(s.get_psi())
        _compare(s, orc, arm, 1.0)


@pytest.mark.parametrize("bc", [A.SMX_BC_NEUMANN, A.SMX_BC_PERIODIC])
def test_surftrack_at_workload_size(bc):
    """The surfTrack workloads' own grids and precision: Neumann 256x256x384 and periodic 256^3,
    FP32-mixed, from the layered start (200 and 20 steps: most columns cross), every column
    against the oracle evaluating the GPU's own field (the evaluation kernels' splits, tables and
    fused z inverse at full size)."""
    dx = 2 * math.pi / 8
    if bc == A.SMX_BC_NEUMANN:
        nx, ny, nz = 256, 256, 384
        cfg = A.make_config(nx, ny, nz, lx=nx * dx, ly=ny * dx, lz=nz * dx, bc=bc, eps=0.5, q0=1.0, beta=0.0,
                            gamma=1.0, dt=0.05, nl_model=A.SMX_NL_CUBIC, precision=A.SMX_FP32_MIXED)
        arm, vsign, steps = 0.7, -1.0, 200
    else:
        nx = ny = nz = 256
        cfg = A.make_config(nx, lx=nx * dx, ly=ny * dx, lz=nz * dx, eps=0.3, q0=1.0, beta=0.0, gamma=1.0, dt=0.1,
                            nl_model=A.SMX_NL_CUBIC, precision=A.SMX_FP32_MIXED)
        arm, vsign, steps = 0.1, 1.0, 20
    c = (np.arange(nx) - nx // 2) * dx
    cz = (np.arange(nz) - nz // 2) * dx
    r2 = np.maximum(c[:, None, None] ** 2 + c[None, :, None] ** 2 - cz[None, None, :] ** 2, 0.0)
    psi0 = 0.4 * np.cos(np.sqrt(r2 + 4.0))
    say("surftrack at workload size", bc)
    with Solver(cfg) as s:
        s.set_psi(psi0)
        s.step(steps)
        orc = O.Oracle(cfg)
        orc.set_psi(s.get_psi())
        (_, _, _, _), found = s.surftrack(arm, vsign)
        assert found > nx * ny // 2
        _compare(s, orc, arm, vsign)
