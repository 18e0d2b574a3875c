"""ctypes mirror of include/smectic.h (constants and the smx_config struct).

Kept in one place so that the host wrapper, the tests and the benchmark agree
on the struct layout.  The layout check against the C compiler lives in the
CPU test suite (tests/test_abi.py).
"""
import ctypes as C

SMX_ABI_VERSION = 1

SMX_BC_PERIODIC = 0
SMX_BC_NEUMANN = 1

SMX_FP64 = 0
SMX_FP32_MIXED = 1

SMX_NL_QUINTIC = 0
SMX_NL_CUBIC = 1

SMX_FORCE_SH = 0
SMX_FORCE_MU = 1

SMX_OK = 0
SMX_ERR_INVALID = -1
SMX_ERR_UNSUPPORTED = -2
SMX_ERR_HIP = -3
SMX_ERR_NONFINITE = -4
SMX_ERR_STATE = -5

ERR_NAMES = {
    SMX_ERR_INVALID: "SMX_ERR_INVALID",
    SMX_ERR_UNSUPPORTED: "SMX_ERR_UNSUPPORTED",
    SMX_ERR_HIP: "SMX_ERR_HIP",
    SMX_ERR_NONFINITE: "SMX_ERR_NONFINITE",
    SMX_ERR_STATE: "SMX_ERR_STATE",
}


class SmxConfig(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int),
        ("bc", C.c_int),
        ("precision", C.c_int),
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("nz", C.c_int),
        ("lx", C.c_double),
        ("ly", C.c_double),
        ("lz", C.c_double),
        ("eps", C.c_double),
        ("q0", C.c_double),
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("gamma", C.c_double),
        ("dt", C.c_double),
        ("nl_model", C.c_int),
        ("advect", C.c_int),
        ("eta", C.c_double),
        ("rho", C.c_double),
        ("force_form", C.c_int),
        ("w_mid", C.c_double),
        ("mid_plane", C.c_int),
        ("device", C.c_int),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class SmxSurfOpts(C.Structure):
    _fields_ = [("arm", C.c_double), ("vel_sign", C.c_double)]


class SmxSchedule(C.Structure):
    """Tile schedules of the periodic step kernels: slots (y inverse, z, y forward, x)."""
    _fields_ = [("nc", C.c_int * 4), ("variant", C.c_int * 4)]

    def as_list(self):
        return [(int(self.nc[k]), int(self.variant[k])) for k in range(4)]

    @classmethod
    def from_list(cls, slots):
        s = cls()
        for k, (nc, v) in enumerate(slots):
            s.nc[k] = nc
            s.variant[k] = v
        return s


def make_config(nx, ny=None, nz=None, *, bc=SMX_BC_PERIODIC, precision=SMX_FP64, lx=None, ly=None,
                lz=None, nw=None, eps=-0.7, q0=1.0, alpha=1.0, beta=2.0, gamma=1.0, dt=1e-3,
                nl_model=SMX_NL_QUINTIC, advect=0, eta=1.0, rho=1.0, force_form=SMX_FORCE_SH,
                w_mid=0.0, mid_plane=-1, device=0):
    """Build an SmxConfig.  Box lengths default to L = n*2*pi/(q0*nw) as in
    periodicFFT/pf3d.cpp:66-71 (nw = points per wavelength, default 16)."""
    import math
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    if nw is None:
        nw = 16
    lx = nx * 2.0 * math.pi / (q0 * nw) if lx is None else lx
    ly = ny * 2.0 * math.pi / (q0 * nw) if ly is None else ly
    lz = nz * 2.0 * math.pi / (q0 * nw) if lz is None else lz
    return SmxConfig(SMX_ABI_VERSION, bc, precision, nx, ny, nz, lx, ly, lz, eps, q0, alpha, beta,
                     gamma, dt, nl_model, advect, eta, rho, force_form, w_mid, mid_plane, device)
