"""ctypes binding of liboracle (oracle/smx_oracle.c) -- TEST INFRASTRUCTURE ONLY.

The oracle is a plain-C FP64 restatement of the reference loops; every
function in smx_oracle.c cites the reference file:line it follows.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from smectic3d_amd._abi import SmxConfig

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsmx_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.POINTER(SmxConfig)]
        L.or_destroy.argtypes = [C.c_void_p]
        for name in ("or_set_psi", "or_get_psi", "or_dtpsi", "or_get_nqpast", "or_velocity"):
            getattr(L, name).argtypes = [C.c_void_p, _dp]
            getattr(L, name).restype = C.c_int
        L.or_step.argtypes = [C.c_void_p, C.c_int64]
        L.or_step.restype = C.c_int
        for name in ("or_l1", "or_energy", "or_energy_fd_trapz", "or_dominant_wavenumber"):
            getattr(L, name).argtypes = [C.c_void_p]
            getattr(L, name).restype = C.c_double
        L.or_surftrack.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                                   C.POINTER(C.c_int64)]
        L.or_surftrack.restype = C.c_int
        L.or_step_count.argtypes = [C.c_void_p]
        L.or_step_count.restype = C.c_int64
        L.or_fft1d.argtypes = [C.c_int, _dp, C.c_int]
        for name in ("or_rfft3", "or_irfft3", "or_redft10_3d", "or_redft01_3d"):
            getattr(L, name).argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
            getattr(L, name).restype = C.c_int
        for name in ("or_redft10_1d", "or_redft01_1d", "or_rodft10_1d", "or_rodft01_1d"):
            getattr(L, name).argtypes = [C.c_int, _dp, _dp]
        L.or_splitmix_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.or_splitmix_uniform.restype = C.c_double
        L.or_ic_noise.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64, C.c_double, _dp]
        L.or_ic_hyperboloid.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_uint64, _dp]
        L.or_growth_sweep.argtypes = [C.POINTER(SmxConfig), C.c_double, C.c_double, C.c_int, _dp]
        L.or_growth_sweep.restype = C.c_int
        L.or_psi_avg.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                 C.c_double, _dp, _dp]
        L.or_psi_avg.restype = C.c_int
        L.or_eckhaus_scan.argtypes = [C.POINTER(SmxConfig), _dp, C.c_int, _dp, _dp, _dp]
        L.or_eckhaus_scan.restype = C.c_int
        L.or_adjoint_grad.argtypes = [C.POINTER(SmxConfig), _dp, C.c_int64, C.c_double, _dp, _dp, _dp, _dp]
        L.or_adjoint_grad.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


# ---- transforms ----
def fft1d(x, sign=-1):
    z = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().or_fft1d(len(z), z.view(np.float64).ctypes.data_as(_dp), sign)
    return z


def rfft3(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    nx, ny, nz = a.shape
    out = np.zeros((nx, ny, nz // 2 + 1), np.complex128)
    lib().or_rfft3(nx, ny, nz, _p(a), out.view(np.float64).ctypes.data_as(_dp))
    return out


def irfft3(c, nz):
    c = np.ascontiguousarray(c, dtype=np.complex128)
    nx, ny, _ = c.shape
    out = np.zeros((nx, ny, nz), np.float64)
    lib().or_irfft3(nx, ny, nz, c.view(np.float64).ctypes.data_as(_dp), _p(out))
    return out


def redft10_3d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros_like(a)
    lib().or_redft10_3d(*a.shape, _p(a), _p(out))
    return out


def redft01_3d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros_like(a)
    lib().or_redft01_3d(*a.shape, _p(a), _p(out))
    return out


def redft10_1d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros_like(a)
    lib().or_redft10_1d(len(a), _p(a), _p(out))
    return out


def redft01_1d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros_like(a)
    lib().or_redft01_1d(len(a), _p(a), _p(out))
    return out


def rodft10_1d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros_like(a)
    lib().or_rodft10_1d(len(a), _p(a), _p(out))
    return out


def rodft01_1d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros_like(a)
    lib().or_rodft01_1d(len(a), _p(a), _p(out))
    return out


# ---- initial conditions ----
def ic_noise(nx, ny, nz, seed, amp, box=0):
    out = np.zeros((nx, ny, nz))
    lib().or_ic_noise(nx, ny, nz, seed, box, amp, _p(out))
    return out


def ic_hyperboloid(nx, ny, nz, dx, dy, dz, q0, a=0.4, noise=1e-3, seed=0):
    out = np.zeros((nx, ny, nz))
    lib().or_ic_hyperboloid(nx, ny, nz, dx, dy, dz, q0, a, noise, seed, _p(out))
    return out


def splitmix_uniform(seed, box, idx):
    return lib().or_splitmix_uniform(seed, box, idx)


def growth_sweep(cfg, mq0, dmq, nq):
    out = np.zeros(nq)
    rc = lib().or_growth_sweep(C.byref(cfg), mq0, dmq, nq, _p(out))
    assert rc == 0
    return out


def eckhaus_scan(cfg, psi0, q, amp):
    """energyPsi.m:68-234: FD + trapz energies of psi0 + amp[m] dpsiz cos((1+q[m]) (k+1) dz)."""
    psi0 = np.ascontiguousarray(psi0, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    amp = np.ascontiguousarray(np.broadcast_to(amp, q.shape), dtype=np.float64)
    out = np.zeros(len(q))
    rc = lib().or_eckhaus_scan(C.byref(cfg), _p(psi0), len(q), _p(q), _p(amp), _p(out))
    assert rc == 0, rc
    return out


def psi_avg(psi, dx, dy, dz, nw, q0=1.0):
    """morphSurfAmp.cpp:1005-1134 literally (O(N^3 nw^3)): small grids only."""
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    out = np.zeros_like(psi)
    rc = lib().or_psi_avg(*psi.shape, dx, dy, dz, nw, q0, _p(psi), _p(out))
    if rc != 0:
        raise ValueError("reference box [nw/2, 3nw/2]^3 does not fit the grid")
    return out


def adjoint_grad(cfg, psi0, T, amp=0.5):
    """FP64 discrete adjoint: returns (J, dJ/dpsi0, dJ/deps, dJ/dq0)."""
    psi0 = np.ascontiguousarray(psi0, dtype=np.float64)
    d = np.zeros_like(psi0)
    J, de, dq = C.c_double(), C.c_double(), C.c_double()
    rc = lib().or_adjoint_grad(C.byref(cfg), _p(psi0), T, amp, C.byref(J), _p(d), C.byref(de), C.byref(dq))
    assert rc == 0, rc
    return J.value, d, de.value, dq.value


def loss_only(cfg, psi0, T, amp=0.5):
    """J = 1/2 sum (psi_T - amp cos(q0 z))^2 from a plain forward run (finite-difference checks)."""
    o = Oracle(cfg)
    o.set_psi(psi0)
    o.step(T)
    p = o.get_psi()
    z = np.arange(cfg.nz) * cfg.lz / cfg.nz
    return 0.5 * float(np.sum((p - amp * np.cos(cfg.q0 * z)) ** 2))


class Oracle:
    """FP64 CPU solver following the reference dataflow (test checker only)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self._s = lib().or_create(C.byref(cfg))
        if not self._s:
            raise ValueError("oracle rejected the configuration")
        self.shape = (cfg.nx, cfg.ny, cfg.nz)

    def __del__(self):
        if getattr(self, "_s", None):
            lib().or_destroy(self._s)
            self._s = None

    def set_psi(self, psi):
        psi = np.ascontiguousarray(psi, dtype=np.float64).reshape(self.shape)
        assert lib().or_set_psi(self._s, _p(psi)) == 0

    def get_psi(self):
        out = np.zeros(self.shape)
        lib().or_get_psi(self._s, _p(out))
        return out

    def nqpast(self):
        nx, ny, nz = self.shape
        if self.cfg.bc == 0:
            out = np.zeros((nx, ny, nz // 2 + 1), np.complex128)
            lib().or_get_nqpast(self._s, out.view(np.float64).ctypes.data_as(_dp))
        else:
            out = np.zeros(self.shape)
            lib().or_get_nqpast(self._s, _p(out))
        return out

    def step(self, n):
        assert lib().or_step(self._s, n) == 0

    def l1(self):
        return lib().or_l1(self._s)

    def energy(self):
        return lib().or_energy(self._s)

    def energy_fd_trapz(self):
        return lib().or_energy_fd_trapz(self._s)

    def dominant_wavenumber(self):
        return lib().or_dominant_wavenumber(self._s)

    def velocity(self):
        """Stokes velocity (3, nx, ny, nz) of the current Neumann state."""
        out = np.zeros((3,) + self.shape)
        assert lib().or_velocity(self._s, _p(out)) == 0
        return out

    def dtpsi(self):
        out = np.zeros(self.shape)
        lib().or_dtpsi(self._s, _p(out))
        return out

    def surftrack(self, arm, vel_sign=1.0, prev=None):
        nx, ny, _ = self.shape
        outs = [np.zeros(nx * ny) if prev is None else prev[i].copy() for i in range(4)]
        found = C.c_int64(0)
        rc = lib().or_surftrack(self._s, arm, vel_sign, *[_p(o) for o in outs], C.byref(found))
        assert rc == 0
        return outs, found.value

    def step_count(self):
        return lib().or_step_count(self._s)
