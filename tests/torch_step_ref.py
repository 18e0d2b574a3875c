"""The periodic CNAB2 step of pf3d.cpp:395-485 (no advection, first N_past from N(psi0),
pf3d.cpp:363-374) restated with torch.fft (rocFFT underneath) in FP64 on the GPU: the
independent FFT library of the cross-check in test_gpu_torch_crosscheck.py.

Run as a separate process (torch's own HIP runtime does not start in a process where
libsmectic's already runs):  python tests/torch_step_ref.py IN.npz OUT.npz
IN holds psi0 and the config scalars; OUT receives psi after nsteps.
"""
import math
import sys

import numpy as np


def torch_steps(c, psi0, nsteps):
    import torch
    dev = torch.device("cuda:0")
    nx, ny, nz = psi0.shape
    kx = 2 * math.pi * torch.fft.fftfreq(nx, d=c["lx"] / nx, dtype=torch.float64, device=dev)
    ky = 2 * math.pi * torch.fft.fftfreq(ny, d=c["ly"] / ny, dtype=torch.float64, device=dev)
    kz = 2 * math.pi * torch.fft.rfftfreq(nz, d=c["lz"] / nz, dtype=torch.float64, device=dev)
    k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    lin = c["eps"] - c["alpha"] * (k2 - c["q0"] ** 2) ** 2
    c1, c2 = 1 + 0.5 * c["dt"] * lin, 1 - 0.5 * c["dt"] * lin

    def nl(p):
        if c["cubic"]:
            return c["beta"] * p * p - c["gamma"] * p ** 3
        return c["beta"] * p ** 3 - c["gamma"] * p ** 5

    psi = torch.tensor(psi0, dtype=torch.float64, device=dev)
    npast = torch.fft.rfftn(nl(psi))
    for _ in range(nsteps):
        ph = torch.fft.rfftn(psi)
        nq = torch.fft.rfftn(nl(psi))
        ph = (c1 * ph + 0.5 * c["dt"] * (3 * nq - npast)) / c2
        npast = nq
        psi = torch.fft.irfftn(ph, s=(nx, ny, nz))
    return psi.cpu().numpy()


def main(argv):
    d = np.load(argv[1])
    c = {k: float(d[k]) for k in ("lx", "ly", "lz", "eps", "alpha", "q0", "beta", "gamma", "dt", "cubic")}
    out = torch_steps(c, d["psi0"], int(d["nsteps"]))
    np.savez(argv[2], psi=out)
    print("torch reference done", out.shape, flush=True)


if __name__ == "__main__":
    main(sys.argv)
