"""Writes tests/golden/transforms.npz: known-answer vectors for the transforms the oracle
restates (FFTW's unnormalised DFT, R2C/C2R, REDFT10/REDFT01), computed with numpy/scipy's
pocketfft on seeded inputs.  Run once in a container with scipy:

    python tests/golden/make_fixtures.py

The reference itself holds no test vectors; these pin the FFTW definitions the oracle
follows (DCT-II y_k = 2 sum x_n cos(pi k (2n+1)/2N) and DCT-III y_k = x_0 + 2 sum_{n>0}
x_n cos(pi n (2k+1)/2N) are scipy's unnormalised types 2 and 3)."""
import os

import numpy as np
import scipy.fft as sf

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "transforms.npz")


def main():
    rng = np.random.default_rng(20240611)
    data = {}
    for n in (12, 16, 24, 30, 48, 96):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        data[f"fft_in_{n}"] = x
        data[f"fft_out_{n}"] = np.fft.fft(x)
        data[f"ifft_out_{n}"] = np.fft.ifft(x) * n  # unnormalised inverse (FFTW BACKWARD)
        r = rng.standard_normal(n)
        data[f"dct_in_{n}"] = r
        data[f"dct2_out_{n}"] = sf.dct(r, type=2)
        data[f"dct3_out_{n}"] = sf.dct(r, type=3)
    a = rng.standard_normal((8, 6, 10))
    data["r2c_in"] = a
    data["r2c_out"] = np.fft.rfftn(a)
    c = np.fft.rfftn(rng.standard_normal((8, 6, 10)))
    data["c2r_in"] = c
    data["c2r_out"] = np.fft.irfftn(c, s=(8, 6, 10), axes=(0, 1, 2)) * (8 * 6 * 10)  # FFTW c2r, unnormalised
    b = rng.standard_normal((6, 8, 10))
    data["redft10_3d_in"] = b
    data["redft10_3d_out"] = sf.dctn(b, type=2)
    data["redft01_3d_out"] = sf.dctn(b, type=3)
    np.savez(OUT, **data)
    print("wrote", OUT, len(data), "arrays")


if __name__ == "__main__":
    main()
