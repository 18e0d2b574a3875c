"""The oracle's transforms against the committed known-answer vectors in
tests/golden/transforms.npz (numpy/scipy pocketfft on seeded inputs; written by
tests/golden/make_fixtures.py).  These pin the FFTW definitions the oracle restates:
unnormalised forward/backward DFT, R2C/C2R, REDFT10 (DCT-II) and REDFT01 (DCT-III)."""
import os

import numpy as np
import pytest

from oracle import oracle as O

FIX = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "transforms.npz"))
LENGTHS = (12, 16, 24, 30, 48, 96)


def close(got, want, tol=1e-12):
    return np.linalg.norm(np.ravel(got - want)) <= tol * np.linalg.norm(np.ravel(want))


@pytest.mark.parametrize("n", LENGTHS)
def test_fft_known_answers(n):
    x = FIX[f"fft_in_{n}"]
    assert close(O.fft1d(x, -1), FIX[f"fft_out_{n}"])
    assert close(O.fft1d(x, +1), FIX[f"ifft_out_{n}"])


@pytest.mark.parametrize("n", LENGTHS)
def test_dct_known_answers(n):
    r = FIX[f"dct_in_{n}"]
    assert close(O.redft10_1d(r), FIX[f"dct2_out_{n}"])
    assert close(O.redft01_1d(r), FIX[f"dct3_out_{n}"])


def test_3d_known_answers():
    assert close(O.rfft3(FIX["r2c_in"]), FIX["r2c_out"])
    assert close(O.irfft3(FIX["c2r_in"], 10), FIX["c2r_out"])
    assert close(O.redft10_3d(FIX["redft10_3d_in"]), FIX["redft10_3d_out"])
    assert close(O.redft01_3d(FIX["redft10_3d_in"]), FIX["redft01_3d_out"])
