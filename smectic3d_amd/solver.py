"""Host-side mirror of the reference time loop over the libsmectic C ABI.

``Solver`` plays the role of one reference driver's state (psi, NqPast, the
operator tables and plans) with the step operator's parameters given at run
time instead of compile time (periodicFFT/pf3d.cpp:35-75).
"""
import ctypes as C

import numpy as np

from . import _abi
from ._lib import check, lib

_dp = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


class Solver:
    def __init__(self, cfg):
        if not isinstance(cfg, _abi.SmxConfig):
            raise TypeError("cfg must be an SmxConfig")
        self.cfg = cfg
        self._h = C.c_void_p()
        check(lib().smx_create(C.byref(cfg), C.byref(self._h)))
        self.shape = (cfg.nx, cfg.ny, cfg.nz)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().smx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- state ----
    def set_psi(self, psi, row_stride=None):
        """psi: (nx, ny, nz) array, or a flat array in the reference layout with row_stride."""
        nz = self.cfg.nz
        if row_stride is None:
            psi = np.ascontiguousarray(psi, dtype=np.float64).reshape(self.shape)
            row_stride = nz
        else:
            psi = np.ascontiguousarray(psi, dtype=np.float64)
        check(lib().smx_set_psi(self._h, _ptr(psi), row_stride))

    def get_psi(self, row_stride=None):
        nx, ny, nz = self.shape
        if row_stride is None:
            out = np.empty(self.shape)
            check(lib().smx_get_psi(self._h, _ptr(out), nz))
        else:
            out = np.zeros(nx * ny * row_stride)
            check(lib().smx_get_psi(self._h, _ptr(out), row_stride))
        return out

    def step(self, n=1):
        check(lib().smx_step(self._h, int(n)))

    def sync(self):
        """Wait for the handle's stream; raises SMX_ERR_NONFINITE if a step check found NaN/Inf."""
        check(lib().smx_sync(self._h))

    def set_l1_tracking(self, on):
        """on: smx_step keeps the state before its last step for l1() (one device copy per call)."""
        check(lib().smx_set_l1_tracking(self._h, 1 if on else 0))

    def set_finite_check(self, every):
        """Steps between the non-finite checks of step() (0: off; default 100)."""
        check(lib().smx_set_finite_check(self._h, int(every)))

    def step_count(self):
        n = C.c_int64()
        check(lib().smx_step_count(self._h, C.byref(n)))
        return n.value

    # ---- diagnostics ----
    def _scalar(self, fn):
        v = C.c_double()
        check(fn(self._h, C.byref(v)))
        return v.value

    def l1(self):
        return self._scalar(lib().smx_l1)

    def energy(self):
        return self._scalar(lib().smx_energy)

    def dominant_wavenumber(self):
        return self._scalar(lib().smx_dominant_wavenumber)

    def dtpsi(self):
        out = np.empty(self.shape)
        check(lib().smx_dtpsi(self._h, _ptr(out), self.cfg.nz))
        return out

    def eckhaus_scan(self, q, amp):
        """Energies of psi0 + amp[m] dpsiz cos((1+q[m])(k+1)dz) for the current state psi0
        (energyPsi.m:68-234); amp may be a scalar."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        amp = np.ascontiguousarray(np.broadcast_to(amp, q.shape), dtype=np.float64)
        out = np.zeros(len(q))
        check(lib().smx_eckhaus_scan(self._h, len(q), _ptr(q), _ptr(amp), _ptr(out)))
        return out

    def psi_avg(self, nw=16):
        """Local amplitude field psiAvg (cosFFT/comet/morphSurfAmp.cpp:1005-1134)."""
        out = np.empty(self.shape)
        check(lib().smx_psi_avg(self._h, int(nw), _ptr(out), self.cfg.nz))
        return out

    def surftrack(self, arm, vel_sign=1.0, out=None):
        """([z index, velocity, H, K], found): each an nx*ny float64 array over the columns, the
        values of the last crossing found in each column (columns without one keep them).  out:
        four such arrays to write in place of four new ones (a caller that tracks every few
        steps skips their allocation and first-touch page faults)."""
        nx, ny, _ = self.shape
        if out is None:
            outs = [np.empty(nx * ny) for _ in range(4)]
        else:
            outs = list(out)
            if len(outs) != 4 or not all(isinstance(o, np.ndarray) and o.dtype == np.float64 and
                                         o.size == nx * ny and o.flags.c_contiguous for o in outs):
                raise ValueError("out: four C-contiguous float64 arrays of nx*ny entries")
        found = C.c_int64()
        opts = _abi.SmxSurfOpts(arm, vel_sign)
        check(lib().smx_surftrack(self._h, C.byref(opts), *[_ptr(o) for o in outs], C.byref(found)))
        return outs, found.value

    def profile_step(self, nsteps):
        ms = np.zeros(16)
        by = np.zeros(16)
        n = C.c_int()
        names = C.create_string_buffer(1024)
        check(lib().smx_profile_step(self._h, int(nsteps), 16, C.byref(n), _ptr(ms), _ptr(by), names, 1024))
        k = n.value
        return list(zip(names.value.decode().split("\n")[:k], ms[:k], by[:k]))

    def adjoint_grad(self, psi0, nsteps, ckpt_every=50, target_amp=0.5):
        """(J, dJ/dpsi0, dJ/deps, dJ/dq0) for J = 1/2 sum (psi_T - target_amp cos(q0 z))^2.
        The handle is left at psi0."""
        psi0 = np.ascontiguousarray(psi0, dtype=np.float64).reshape(self.shape)
        d = np.empty(self.shape)
        J, de, dq = C.c_double(), C.c_double(), C.c_double()
        check(lib().smx_adjoint_grad(self._h, _ptr(psi0), self.cfg.nz, int(nsteps), int(ckpt_every),
                                     float(target_amp), C.byref(J), _ptr(d), C.byref(de), C.byref(dq)))
        return J.value, d, de.value, dq.value

    def set_adjoint_budget(self, nbytes):
        """Device bytes of resident adjoint records: -1 (default) lends up to 80% of the free
        memory for one call; >= 0 keeps a workspace of at most that many bytes between calls
        (0: checkpoints only)."""
        check(lib().smx_set_adjoint_budget(self._h, int(nbytes)))

    def adjoint_info(self):
        """(resident steps, recomputed steps, device bytes) of the last adjoint_grad."""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().smx_adjoint_info(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def save_checkpoint(self, path):
        """Exact binary checkpoint (carried state, FP64 mean mode, step count)."""
        check(lib().smx_save_checkpoint(self._h, str(path).encode()))

    def load_checkpoint(self, path):
        check(lib().smx_load_checkpoint(self._h, str(path).encode()))

    def calibration_copy(self, nbytes, reps=20):
        v = C.c_double()
        check(lib().smx_calibration_copy(self._h, int(nbytes), int(reps), C.byref(v)))
        return v.value

    def schedule(self):
        """[(nc, variant)] of the step kernels (y inverse, z, y forward, x)."""
        s = _abi.SmxSchedule()
        check(lib().smx_get_schedule(self._h, C.byref(s)))
        return s.as_list()

    def set_schedule(self, slots):
        check(lib().smx_set_schedule(self._h, C.byref(_abi.SmxSchedule.from_list(slots))))

    def autotune(self, reps=20):
        """Time every candidate tile schedule and install the fastest (state untouched)."""
        s = _abi.SmxSchedule()
        check(lib().smx_autotune(self._h, int(reps), C.byref(s)))
        return s.as_list()


class Batch:
    """Independent boxes of uneven sizes advanced together through one captured hipGraph."""

    def __init__(self, cfgs):
        arr = (_abi.SmxConfig * len(cfgs))(*cfgs)
        self._b = C.c_void_p()
        check(lib().smx_batch_create(len(cfgs), arr, C.byref(self._b)))
        self.cfgs = list(cfgs)
        self.boxes = []
        for i, c in enumerate(cfgs):
            h = C.c_void_p()
            check(lib().smx_batch_handle(self._b, i, C.byref(h)))
            s = Solver.__new__(Solver)
            s.cfg = c
            s._h = h
            s.shape = (c.nx, c.ny, c.nz)
            s.close = lambda: None  # owned by the batch
            self.boxes.append(s)

    def run(self, nsteps, graph_steps=10):
        check(lib().smx_batch_run(self._b, int(nsteps), int(graph_steps)))

    def autotune(self, reps=20):
        """Autotune every box's tile schedules (one tuning per box shape); the next run
        captures the graph again."""
        check(lib().smx_batch_autotune(self._b, int(reps)))
        return [b.schedule() for b in self.boxes]

    def graph_nodes(self):
        n = C.c_int64()
        check(lib().smx_batch_graph_nodes(self._b, C.byref(n)))
        return n.value

    def adjoint_grad(self, psi0s, nsteps, ckpt_every=50, target_amps=0.5):
        """Gradients of every box in one captured graph: lists (J, dJ/dpsi0, dJ/deps, dJ/dq0),
        J_i = 1/2 sum (psi_T - target_amps[i] cos(q0_i z))^2.  Every box is left at its psi0."""
        nb = len(self.boxes)
        psi0s = [np.ascontiguousarray(p, dtype=np.float64).reshape(b.shape) for p, b in zip(psi0s, self.boxes)]
        if len(psi0s) != nb:
            raise ValueError("one psi0 per box")
        nz = {b.cfg.nz for b in self.boxes}
        stride = max(nz)
        if len(nz) > 1:  # one row stride for every box: pad the rows of the shorter boxes
            ins = []
            for p in psi0s:
                q = np.zeros(p.shape[:2] + (stride,))
                q[..., :p.shape[2]] = p
                ins.append(q)
        else:
            ins = psi0s
        outs = [np.zeros(p.shape[:2] + (stride,)) for p in psi0s]
        amps = np.ascontiguousarray(np.broadcast_to(np.asarray(target_amps, dtype=np.float64), (nb,)))
        J, de, dq = np.zeros(nb), np.zeros(nb), np.zeros(nb)
        pin = (_dp * nb)(*[_ptr(p) for p in ins])
        pout = (_dp * nb)(*[_ptr(o) for o in outs])
        check(lib().smx_batch_adjoint_grad(self._b, pin, stride, int(nsteps), int(ckpt_every), _ptr(amps), _ptr(J),
                                           pout, _ptr(de), _ptr(dq)))
        grads = [o[..., :p.shape[2]].copy() for o, p in zip(outs, psi0s)]
        return list(J), grads, list(de), list(dq)

    def adjoint_nodes(self):
        n = C.c_int64()
        check(lib().smx_batch_adjoint_nodes(self._b, C.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "_b", None) and self._b.value:
            for s in self.boxes:
                s._h = C.c_void_p()
            lib().smx_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
