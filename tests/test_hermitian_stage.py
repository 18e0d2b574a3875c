"""Why FP64 copies of a periodic FP32 stage are projected onto Hermitian symmetry on the kz = 0
and kz = nz/2 planes (k_herm_stage): a numpy model of the stage ([kz][x][ky], x real) with
anti-Hermitian noise on those planes.  The C2R of the field drops the noise, and so do even
multipliers; a first derivative with the +pi/d Nyquist wavenumber (pf3d.cpp:162-190) does not,
unless the stage is projected first (G(x, ky) <- (G(x, ky) + conj G(x, -ky)) / 2)."""
import numpy as np

NX, NY, NZ, DX = 16, 24, 12, 0.7


def _k(n):
    k = np.fft.fftfreq(n, d=DX) * 2 * np.pi
    k[n // 2] = np.pi / DX
    return k


def _stage_with_noise(seed):
    rng = np.random.default_rng(seed)
    psi = rng.standard_normal((NX, NY, NZ))
    G = np.fft.fft(np.fft.rfft(psi, axis=2), axis=1)
    for kz in (0, NZ // 2):
        N = np.fft.fft(rng.standard_normal((NX, NY)) + 1j * rng.standard_normal((NX, NY)), axis=1) * 1e-3
        G[:, :, kz] += 0.5 * (N - np.conj(N[:, (-np.arange(NY)) % NY]))
    return G


def _herm(G):
    G = G.copy()
    for kz in (0, NZ // 2):
        a = G[:, :, kz]
        G[:, :, kz] = 0.5 * (a + np.conj(a[:, (-np.arange(NY)) % NY]))
    return G


def _from_stage(G, mult):  # the diagnostic chain: x forward, multiplier, x and y inverse, z C2R
    P = np.fft.fft(G, axis=0) * mult
    return np.fft.irfft(np.fft.ifft(np.fft.ifft(P, axis=0), axis=1), n=NZ, axis=2)


def _from_field(G, mult):  # the oracle: the same derivative of the real field the stage stands for
    psi = np.fft.irfft(np.fft.ifft(G, axis=1), n=NZ, axis=2)
    P = np.fft.fft(np.fft.fft(np.fft.rfft(psi, axis=2), axis=1), axis=0) * mult
    return np.fft.irfft(np.fft.ifft(np.fft.ifft(P, axis=0), axis=1), n=NZ, axis=2)


def test_projection_makes_stage_derivatives_those_of_the_field():
    G = _stage_with_noise(3)
    kx = _k(NX)[:, None, None] * np.ones((1, NY, NZ // 2 + 1))
    ky = _k(NY)[None, :, None] * np.ones((NX, 1, NZ // 2 + 1))
    for name, m, odd in (("i kx", 1j * kx, True), ("i ky", 1j * ky, True), ("-kx^2", -kx ** 2, False)):
        raw = np.abs(_from_stage(G, m) - _from_field(G, m)).max()
        proj = np.abs(_from_stage(_herm(G), m) - _from_field(G, m)).max()
        assert proj < 1e-12, name
        assert (raw > 1e-6) == odd, name  # only the odd derivatives see the noise
    # the field itself is unchanged by the projection
    assert np.abs(_from_stage(_herm(G), 1.0) - _from_stage(G, 1.0)).max() < 1e-12
