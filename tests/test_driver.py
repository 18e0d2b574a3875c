"""The C++ driver (smectic3d_amd/driver/smx_run): initial conditions and the
reference text-file format, checked on the CPU (--ic-only needs no device)."""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from smectic3d_amd import build as B


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(B.DRIVER):
        B.build_driver()
    return B.DRIVER


def run_ic(driver, tmp_path, *args):
    subprocess.run([driver, "--ic-only", "--out", str(tmp_path), *args], check=True)
    return (tmp_path / "psiIC.dat").read_text()


def test_dump_format_matches_reference_writer(driver, tmp_path):
    """`os << v << "\\n "` per value (pf3d.cpp:351): default ostream precision, 6 significant digits."""
    txt = run_ic(driver, tmp_path, "--preset", "plain", "--n", "16", "--seed", "5", "--amp", "0.01")
    lines = txt.split("\n ")
    assert lines[-1] == "" and len(lines) == 16 ** 3 + 1
    vals = np.array([float(v) for v in lines[:-1]])
    want = O.ic_noise(16, 16, 16, 5, 0.01).ravel()
    for s, w in zip(lines[:50], want[:50]):
        assert s == "%.6g" % w
    assert np.abs(vals - want).max() <= 5e-6 * np.abs(want).max()


def test_hyperboloid_ic_matches_oracle(driver, tmp_path):
    txt = run_ic(driver, tmp_path, "--preset", "hyper", "--n", "16", "--amp", "1e-3", "--seed", "0")
    vals = np.array([float(v) for v in txt.split("\n ")[:-1]]).reshape(16, 16, 16)
    dx = 2 * math.pi / 8
    want = O.ic_hyperboloid(16, 16, 16, dx, dx, dx, 1.0, 0.4, 1e-3, 0)
    np.testing.assert_allclose(vals, want, rtol=1e-5, atol=1e-6)


def test_file_ic_rejects_short_file(driver, tmp_path):
    p = tmp_path / "short.dat"
    p.write_text("0.1\n 0.2\n ")
    r = subprocess.run([driver, "--ic-only", "--out", str(tmp_path), "--n", "16", "--ic", f"file:{p}"],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "short" in r.stderr


@pytest.mark.gpu
def test_driver_resume_from_checkpoint_is_exact(driver, tmp_path):
    """100 steps, --checkpoint, then --resume for 100 more: the final psi.dat equals that of
    one 200-step run (the reference's text restart re-bootstraps and is not exact)."""
    common = ["--preset", "plain", "--n", "32", "--l1-every", "50", "--save-every", "1000"]
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    ck = tmp_path / "state.ckpt"
    subprocess.run([driver, *common, "--steps", "200", "--out", str(a)], check=True)
    subprocess.run([driver, *common, "--steps", "100", "--out", str(b), "--checkpoint", str(ck)], check=True)
    subprocess.run([driver, *common, "--steps", "100", "--out", str(b), "--resume", str(ck)], check=True)
    assert (a / "psi.dat").read_text() == (b / "psi.dat").read_text()


@pytest.mark.gpu
def test_driver_psimid_slice(driver, tmp_path):
    """psiMid.dat gets the (x, z) slice at y = ny/2 at each surface output, empty file at start
    (cosNoAdvMorphCurv.cpp:482-486, 981-1001); the last one equals that slice of the final psi.dat."""
    nx, ny, nz = 16, 32, 24
    subprocess.run([driver, "--preset", "plain", "--nx", str(nx), "--ny", str(ny), "--nz", str(nz), "--neumann",
                    "--l1-every", "10", "--surf-every", "2", "--steps", "40", "--save-every", "1000",
                    "--out", str(tmp_path)], check=True)
    mid = np.array([float(v) for v in (tmp_path / "psiMid.dat").read_text().split()])
    assert mid.size == 2 * nx * nz  # two surface outputs (after 20 and 40 steps)
    psi = np.array([float(v) for v in (tmp_path / "psi.dat").read_text().split()]).reshape(nx, ny, nz)
    np.testing.assert_allclose(mid[nx * nz:].reshape(nx, nz), psi[:, ny // 2, :], rtol=0, atol=0)


@pytest.mark.gpu
def test_driver_psi_avg_file(driver, tmp_path):
    """--psi-avg writes psiAvg.dat of the final state in the dump format; its values are the
    oracle's psiAvg of the printed psi.dat to the 6 printed digits."""
    n, nw = 16, 4
    subprocess.run([driver, "--preset", "plain", "--n", str(n), "--l1-every", "20", "--steps", "40",
                    "--save-every", "1000", "--psi-avg", str(nw), "--out", str(tmp_path)], check=True)
    txt = (tmp_path / "psiAvg.dat").read_text()
    assert txt.endswith("\n ")
    got = np.array([float(v) for v in txt.split()]).reshape(n, n, n)
    psi = np.array([float(v) for v in (tmp_path / "psi.dat").read_text().split()]).reshape(n, n, n)
    d = 2 * math.pi / 8  # plain preset: 8 points per wavelength, q0 = 1
    want = O.psi_avg(psi, d, d, d, nw)
    np.testing.assert_allclose(got, want, rtol=1e-4)


@pytest.mark.gpu
@pytest.mark.parametrize("bc", ["periodic", "neumann"])
def test_driver_advected_run_with_surfaces(driver, tmp_path, bc):
    """The advected loop with surface outputs (surfTrack.cpp is an advected driver): the surf
    files get one record per surface output and velSurf.dat carries dpsi/dt with advection."""
    nx, ny, nz = 16, 16, 32
    extra = ["--neumann", "--advect", "1", "--arm", "0.7"] if bc == "neumann" else []
    subprocess.run([driver, "--preset", "advected", "--nx", str(nx), "--ny", str(ny), "--nz", str(nz), *extra,
                    "--l1-every", "10", "--surf-every", "2", "--steps", "40", "--save-every", "1000",
                    "--out", str(tmp_path)], check=True)
    for name in ("surfPsi.dat", "velSurf.dat", "curvH.dat", "curvK.dat"):
        assert (tmp_path / name).exists()
    vel = np.array([float(v) for v in (tmp_path / "velSurf.dat").read_text().split()])
    assert vel.size > 0 and np.all(np.isfinite(vel))
