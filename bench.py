"""Benchmark of the periodic Swift-Hohenberg step on one MI355X.

Headline workload (pf3dMorph-style periodic run): 256^3 grid, L = 32*2*pi per axis
(8 points per wavelength), q0 = 1, eps = 0.3, beta = 0, gamma = 1, dt = 0.1, cubic
nonlinearity, CNAB2; psi0 = 0.4 cos(q0 sqrt(max(x^2+y^2-z^2, 0) + 4)) about the box centre
plus uniform noise of amplitude 1e-3 (SplitMix64 seed 0).  `value` is the FP64 throughput
(the reference's arithmetic); the FP32-mixed path runs beside it from the same state, and
its error against the FP64 path after the workload's full 2000 steps is stated and checked
against the 1e-6 / 1e-9 bounds.  Secondary legs: the uneven 24-box batch (forward graph and
batched gradients), the Neumann box with surfTrack histograms, the advected boxes (with
their FP32-vs-FP64 error) and the 256^3 adjoint.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SH steps/s at 256³ (FP64 vs FP32-mixed, ≤1e-6 rel); adjoint grads/s; %HBM peak"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_cfg(precision, n=256):
    from smectic3d_amd import _abi as A
    L = n / 8 * 2 * math.pi
    return A.make_config(n, lx=L, ly=L, lz=L, eps=0.3, q0=1.0, beta=0.0, gamma=1.0, dt=0.1,
                         nl_model=A.SMX_NL_CUBIC, precision=precision)


def hyperboloid_ic(n, dx, q0=1.0, a=0.4, noise=1e-3, seed=0):
    """psi0 = a cos(q0 sqrt(max(x^2+y^2-z^2,0)+4)) + U[-noise, noise] (SplitMix64, counter = linear index)."""
    c = (np.arange(n) - n // 2) * dx
    x = c[:, None, None]
    y = c[None, :, None]
    z = c[None, None, :]
    r2 = np.maximum(x * x + y * y - z * z, 0.0)
    psi = a * np.cos(q0 * np.sqrt(r2 + 4.0))
    idx = np.arange(n * n * n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        zz = np.uint64(seed) + (idx + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        zz = (zz ^ (zz >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        zz = (zz ^ (zz >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        zz = zz ^ (zz >> np.uint64(31))
    u = (zz >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return psi + noise * (2.0 * u.reshape(n, n, n) - 1.0)


SCHED_FILE = os.path.join(ROOT, "profiles", "periodic_256_schedules.json")


def pinned_schedules():
    """Tile schedules of the 256^3 step kernels recorded by an autotuned run and used by the
    profiled runs under profiles/ (so the kernels timed here are the ones profiled)."""
    try:
        with open(SCHED_FILE) as f:
            d = json.load(f)
        return d["fp32"], d["fp64"]
    except (OSError, ValueError, KeyError):
        return None, None


def run_path(precision, psi0, steps, warmup, prof_steps, tune=True, fixed=None):
    """Timed region: `steps` steps in one smx_step call on the resident state, L1 tracking off
    (no snapshot copy inside the region), bracketed by stream synchronisations."""
    from smectic3d_amd import Solver
    cfg = workload_cfg(precision)
    s = Solver(cfg)
    s.set_psi(psi0)
    s.set_l1_tracking(False)
    # planning, as FFTW_MEASURE at set-up (pf3d.cpp:226-291): the autotuner times every tile
    # schedule on a scratch copy of the loaded state and installs the fastest
    t0 = time.perf_counter()
    if fixed is not None:  # given schedules (profiled runs of fixed kernels)
        s.set_schedule([tuple(x) for x in fixed])
        sched = s.schedule()
    else:
        sched = s.autotune(20) if tune else s.schedule()
    tune_s = time.perf_counter() - t0
    s.step(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    t1 = time.perf_counter()
    out = {"seconds": t1 - t0, "steps_per_s": steps / (t1 - t0), "ms_per_step": 1e3 * (t1 - t0) / steps,
           "schedule": sched, "tune_s": tune_s}
    out["kernels"] = s.profile_step(prof_steps)
    if precision == 1:
        # plain float4 copy of one stage array: the achievable-bandwidth yardstick
        n_spec_bytes = cfg.nx * cfg.ny * (cfg.nz // 2 + 1) * 8
        out["copy_GBs"] = 2 * n_spec_bytes / (s.calibration_copy(n_spec_bytes, 50) * 1e-3) / 1e9
    s.close()
    return out


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def run_accuracy(psi0, nsteps=2000, f32=None, f64=None):
    """FP32-mixed against FP64 over the workload's full 2000 steps (independent of --steps):
    relative L2 of psi, relative energy error and relative error of the dominant wavenumber,
    checked against the 1e-6 (psi) and 1e-9 (energy) bounds."""
    from smectic3d_amd import Solver
    from smectic3d_amd import _abi as A
    res = {}
    for prec, sch in ((A.SMX_FP64, f64), (A.SMX_FP32_MIXED, f32)):
        with Solver(workload_cfg(prec)) as s:
            if sch is not None:
                s.set_schedule([tuple(x) for x in sch])
            s.set_psi(psi0)
            s.step(nsteps)
            res[prec] = (s.get_psi(), s.energy(), s.dominant_wavenumber(), s.l1())
    p64, e64, q64, l64 = res[A.SMX_FP64]
    p32, e32, q32, l32 = res[A.SMX_FP32_MIXED]
    out = {"steps": nsteps, "psi_rel_l2": rel(p32, p64), "energy_rel": abs(e32 - e64) / abs(e64),
           "qdom_rel": abs(q32 - q64) / abs(q64), "energy_fp64": e64, "qdom_fp64": q64, "l1_fp64": l64,
           "l1_fp32_mixed": l32}
    out["meets_psi_1e-6"] = out["psi_rel_l2"] <= 1e-6
    out["meets_energy_1e-9"] = out["energy_rel"] <= 1e-9
    return out


def neumann_cfg(precision):
    from smectic3d_amd import _abi as A
    nx, ny, nz = 256, 256, 384
    dx = 2 * math.pi / 8
    return A.make_config(nx, ny, nz, lx=nx * dx, ly=ny * dx, lz=nz * dx, bc=A.SMX_BC_NEUMANN, eps=0.5, q0=1.0,
                         beta=0.0, gamma=1.0, dt=0.05, nl_model=A.SMX_NL_CUBIC, precision=precision)


def histogram_compare(a, b, bins=64):
    """Histograms of the FP32 values `a` and the FP64 values `b` on common bins: 64 equal bins
    between the 0.5th and 99.5th percentiles of the FP64 values (values outside go to the end
    bins).  Returns the total-variation distance of the two normalised histograms and the
    relative differences of mean and standard deviation."""
    lo, hi = np.percentile(b, [0.5, 99.5])
    edges = np.linspace(lo, hi, bins + 1)
    ha = np.histogram(np.clip(a, lo, hi), edges)[0] / max(1, len(a))
    hb = np.histogram(np.clip(b, lo, hi), edges)[0] / max(1, len(b))
    return {"bins": bins, "range_fp64_p0.5_p99.5": [float(lo), float(hi)], "samples": [int(len(a)), int(len(b))],
            "total_variation": float(0.5 * np.abs(ha - hb).sum()),
            "mean_rel": float(abs(a.mean() - b.mean()) / abs(b.mean())),
            "std_rel": float(abs(a.std() - b.std()) / b.std())}


def run_neumann(steps, surf_every=20, compare=True):
    """cosNoAdvMorph DCT/Neumann + surfTrack: 256x256x384, FP32 bulk with FP64 reductions,
    eps=0.5, q0=1, gamma=1, dt=0.05, hyperboloid IC; surfTrack every 20 steps (the reference's
    column scan from the top, arm 0.7, v = -dpsi/dt / |grad psi|).  The same run in FP64 gives
    the reference histograms of v and H over every call's crossing columns."""
    from smectic3d_amd import Solver
    from smectic3d_amd import _abi as A
    nx, ny, nz = 256, 256, 384
    dx = 2 * math.pi / 8
    c = (np.arange(nx) - nx // 2) * dx
    cz = (np.arange(nz) - nz // 2) * dx
    r2 = np.maximum(c[:, None, None] ** 2 + c[None, :, None] ** 2 - cz[None, None, :] ** 2, 0.0)
    psi0 = 0.4 * np.cos(np.sqrt(r2 + 4.0))
    out, samples = {}, {}
    ncalls, nxy = steps // surf_every, nx * ny
    for prec in ((A.SMX_FP32_MIXED, A.SMX_FP64) if compare else (A.SMX_FP32_MIXED,)):
        # output arrays written in place, touched before the timed loop: with compare, call i's
        # velocity and H go straight into row i of the sample store (every column reported on that
        # call, the persistent arrays as the reference's)
        bufs = np.empty((4, nxy))
        keep = np.empty((2, ncalls, nxy)) if compare else None
        for b in (bufs, keep):
            if b is not None:
                b.fill(0.0)
        with Solver(neumann_cfg(prec)) as s:
            s.set_psi(psi0)
            s.set_l1_tracking(False)
            sched = s.autotune(10)  # tile widths of the Neumann step kernels
            s.sync()
            t0 = time.perf_counter()
            found = 0
            for i in range(ncalls):
                s.step(surf_every)
                o = (bufs[0], keep[0, i], keep[1, i], bufs[3]) if compare else bufs
                _, f = s.surftrack(0.7, -1.0, out=o)
                found += f
            s.sync()
            t1 = time.perf_counter()
            psi_end = s.get_psi() if compare else None
            if prec == A.SMX_FP32_MIXED:
                ks = s.profile_step(20)
                t_s0 = time.perf_counter()
                for _ in range(20):
                    s.surftrack(0.7, -1.0, out=bufs)
                t_surf = (time.perf_counter() - t_s0) / 20
        samples[prec] = (keep[0].ravel() if compare else None, keep[1].ravel() if compare else None, psi_end)
        key = "fp32" if prec == A.SMX_FP32_MIXED else "fp64"
        out[key] = {"steps_per_s_with_surftrack": steps / (t1 - t0),
                    "columns_found_per_call": found / max(1, steps // surf_every), "schedule": sched}
    r = out["fp32"]
    # the calls' share of the run: the separately timed call cost times the calls, over the run
    run_s = steps / r["steps_per_s_with_surftrack"]
    r.update({"steps": steps, "surftrack_every": surf_every, "surftrack_ms_per_call": 1e3 * t_surf,
              "surftrack_share_of_run": ncalls * t_surf / run_s,
              "kernels": [{"name": k, "avg_ms": m, "GBs": b / (m * 1e-3) / 1e9} for k, m, b in ks]})
    if compare:
        v32, h32, p32 = samples[A.SMX_FP32_MIXED]
        v64, h64, p64 = samples[A.SMX_FP64]
        r["fp64"] = out["fp64"]
        r["vs_fp64"] = {"psi_rel_l2_end": rel(p32, p64),
                        "velocity_hist": histogram_compare(v32, v64), "meancurv_hist": histogram_compare(h32, h64),
                        "note": "histograms over every column's value at each of the surfTrack calls "
                                "(the persistent output arrays), FP32-mixed run against the FP64 run"}
    return r


def advected_cfg(precision):
    from smectic3d_amd import _abi as A
    n = 256
    L = n / 8 * 2 * math.pi
    return A.make_config(n, lx=L, ly=L, lz=L, eps=0.5, q0=1.0, beta=0.0, gamma=1.0, dt=0.02,
                         nl_model=A.SMX_NL_CUBIC, advect=1, eta=1.0, rho=1.0, force_form=A.SMX_FORCE_MU,
                         precision=precision)


def run_surftrack_periodic(calls=20):
    """Periodic surfTrack at 256^3 FP32-mixed, without advection (the field of the headline
    workload after 20 steps) and with it (the advected workload's noise start after 20 steps):
    milliseconds per call, the outputs written into caller-owned arrays."""
    from smectic3d_amd import Solver
    from smectic3d_amd import _abi as A
    n = 256
    bufs = np.empty((4, n * n))
    bufs.fill(0.0)
    out = {}
    for name, cfg, psi0, arm in (
            ("plain", workload_cfg(A.SMX_FP32_MIXED), hyperboloid_ic(n, 2 * math.pi / 8), 0.1),
            ("advected", advected_cfg(A.SMX_FP32_MIXED), np.random.default_rng(4).uniform(-0.05, 0.05, (n, n, n)),
             0.02)):
        with Solver(cfg) as s:
            s.set_psi(psi0)
            s.set_l1_tracking(False)
            s.step(20)
            s.sync()
            for _ in range(3):
                s.surftrack(arm, 1.0, out=bufs)
            t0 = time.perf_counter()
            found = 0
            for _ in range(calls):
                found += s.surftrack(arm, 1.0, out=bufs)[1]
            out[name] = {"ms_per_call": 1e3 * (time.perf_counter() - t0) / calls, "columns_found_per_call": found / calls}
    return {"grid": [n, n, n], "precision": "fp32_mixed", "calls": calls, **out}


def run_advected(steps, compare=True):
    """Advection-coupled periodic 256^3: mu grad psi force, spectral Stokes (eta=1), eps=0.5, q0=1,
    gamma=1, dt=0.02, psi0 ~ U[-0.05, 0.05]; FP32-mixed, with the same run in FP64 for its error."""
    from smectic3d_amd import Solver
    from smectic3d_amd import _abi as A
    n = 256
    psi0 = np.random.default_rng(4).uniform(-0.05, 0.05, (n, n, n))
    with Solver(advected_cfg(A.SMX_FP32_MIXED)) as s:
        s.autotune(10)
        s.set_psi(psi0)
        s.set_l1_tracking(False)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        t1 = time.perf_counter()
        e = s.energy()
        p32 = s.get_psi() if compare else None
        ks = s.profile_step(10)
    out = {"steps": steps, "steps_per_s": steps / (t1 - t0), "energy": e,
           "kernels": [{"name": k, "avg_ms": m, "GBs": b / (m * 1e-3) / 1e9} for k, m, b in ks]}
    if compare:
        with Solver(advected_cfg(A.SMX_FP64)) as s:
            s.set_psi(psi0)
            s.set_l1_tracking(False)
            s.sync()
            t0 = time.perf_counter()
            s.step(steps)
            s.sync()
            t1 = time.perf_counter()
            e64 = s.energy()
            p64 = s.get_psi()
        out["fp64_steps_per_s"] = steps / (t1 - t0)
        out["vs_fp64"] = {"psi_rel_l2": rel(p32, p64), "energy_rel": abs(e - e64) / abs(e64), "energy_fp64": e64}
    return out


def run_neumann_advected(steps):
    """Advection-coupled Neumann box 256x256x384: the DCT grid of cosNoAdvMorph with the
    sine/cosine Stokes projection and the advected workload's parameters (mu grad psi force,
    eta=1, eps=0.5, q0=1, gamma=1, dt=0.02, psi0 ~ U[-0.05, 0.05]), 8 points per wavelength;
    FP32-mixed.  (A 0.4-amplitude layered start is unstable at this dt under explicit advection,
    in the FP64 oracle as on the GPU.)"""
    from smectic3d_amd import Solver
    from smectic3d_amd import _abi as A
    nx, ny, nz = 256, 256, 384
    dx = 2 * math.pi / 8
    cfg = A.make_config(nx, ny, nz, lx=nx * dx, ly=ny * dx, lz=nz * dx, bc=A.SMX_BC_NEUMANN, eps=0.5, q0=1.0,
                        beta=0.0, gamma=1.0, dt=0.02, nl_model=A.SMX_NL_CUBIC, advect=1, eta=1.0, rho=1.0,
                        force_form=A.SMX_FORCE_MU, precision=A.SMX_FP32_MIXED)
    rng = np.random.default_rng(5)
    with Solver(cfg) as s:
        s.set_psi(rng.uniform(-0.05, 0.05, (nx, ny, nz)))
        s.step(5)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        t1 = time.perf_counter()
        e = s.energy()
        ks = s.profile_step(10)
    S = nx * ny * nz * 4  # one real FP32 stage array
    sps = steps / (t1 - t0)
    alg = 20 * S  # minimal DAG of the advected step, 8 S_r + 12 S_c, with S_r = S_c = S on the DCT grid
    return {"steps": steps, "steps_per_s": sps, "energy": e,
            "step_roofline": {"alg_bytes_per_step": alg, "achieved_GBs": alg * sps / 1e9,
                              "frac": alg * sps / 1e9 / HBM_PEAK_GBS},
            "kernels": [{"name": k, "avg_ms": m, "GBs": b / (m * 1e-3) / 1e9} for k, m, b in ks]}


def batch_cfgs():
    """24 independent periodic boxes, sizes from {64^3, 96^3, 128^3} by seed, eps ~ U[0.2,0.6],
    q0 ~ U[0.8,1.2], gamma=1, dt=0.05 (FP32-mixed)."""
    from smectic3d_amd import _abi as A
    rng = np.random.default_rng(2024)
    cfgs = []
    for _ in range(24):
        n = int(rng.choice([64, 96, 128]))
        L = n / 8 * 2 * math.pi
        cfgs.append(A.make_config(n, lx=L, ly=L, lz=L, eps=float(rng.uniform(0.2, 0.6)),
                                  q0=float(rng.uniform(0.8, 1.2)), beta=0.0, gamma=1.0, dt=0.05,
                                  nl_model=A.SMX_NL_CUBIC, precision=A.SMX_FP32_MIXED))
    return cfgs


def run_batch(steps, tune=True):
    """The 24 boxes of batch_cfgs(), noise 0.05, advanced in one captured hipGraph."""
    from smectic3d_amd import Batch
    cfgs = batch_cfgs()
    pts = sum(c.nx * c.ny * c.nz for c in cfgs)
    with Batch(cfgs) as b:
        t0 = time.perf_counter()
        scheds = b.autotune(10) if tune else [s.schedule() for s in b.boxes]
        tune_s = time.perf_counter() - t0
        for i, (s, c) in enumerate(zip(b.boxes, cfgs)):
            s.set_psi(np.random.default_rng(i).uniform(-0.05, 0.05, (c.nx, c.ny, c.nz)))
        b.run(10, graph_steps=10)
        t0 = time.perf_counter()
        b.run(steps, graph_steps=10)
        t1 = time.perf_counter()
        nodes = b.graph_nodes()
    return {"boxes": 24, "sizes": sorted(c.nx for c in cfgs), "steps": steps,
            "batch_steps_per_s": steps / (t1 - t0), "box_steps_per_s": 24 * steps / (t1 - t0),
            "points_per_s": pts * steps / (t1 - t0), "graph_nodes": nodes, "tune_s": tune_s,
            "schedules": {f"{c.nx}^3": sc for c, sc in zip(cfgs, scheds)}}


def run_batch_adjoint(steps=500, ckpt=50, reps=3):
    """Batched gradients of the 24 uneven boxes of batch_cfgs() (noise 0.05): loss
    1/2 |psi_T - 0.5 cos(q0 z)|^2 accumulated in FP64, dJ/dpsi0, dJ/deps, dJ/dq0 by the
    adjoint with checkpoints every 50 steps, every box's sweep a branch of one captured
    hipGraph launched once per call.  Each box keeps its workspace (an explicit budget of
    its full records), so repeated calls reuse the captured graph; the first call's time
    includes capturing and instantiating it.  Compared with the same boxes' forward solve
    as one graph batch (host to host: psi0 in, psi_T out)."""
    from smectic3d_amd import Batch
    cfgs = batch_cfgs()
    psis = [np.random.default_rng(100 + i).uniform(-0.05, 0.05, (c.nx, c.ny, c.nz)) for i, c in enumerate(cfgs)]
    with Batch(cfgs) as b:
        b.autotune(10)
        for s, c in zip(b.boxes, cfgs):
            n = c.nx * c.ny * c.nz
            sbytes = 8 * c.nx * c.ny * (c.nz // 2 + 1)
            s.set_adjoint_budget(((steps + ckpt - 1) // ckpt) * (4 * n * ckpt + sbytes * (ckpt + 1)))
        t0 = time.perf_counter()
        J, d, de, dq = b.adjoint_grad(psis, steps, ckpt, 0.5)
        tfirst = time.perf_counter() - t0
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            J2, d2, de2, dq2 = b.adjoint_grad(psis, steps, ckpt, 0.5)
            ts.append(time.perf_counter() - t0)
        same = all(np.array_equal(x, y) for x, y in zip(d, d2)) and de == de2 and dq == dq2
        nodes = b.adjoint_nodes()
        resident = [s.adjoint_info() for s in b.boxes]
        # forward solve of the same boxes and length, one graph batch, host to host
        t0 = time.perf_counter()
        for s, p in zip(b.boxes, psis):
            s.set_psi(p)
        b.run(steps, graph_steps=50)
        for s in b.boxes:
            s.get_psi()
        tfwd = time.perf_counter() - t0
    tg = float(np.median(ts))
    return {"boxes": len(cfgs), "sizes": sorted(c.nx for c in cfgs), "steps": steps, "ckpt_every": ckpt,
            "grads_per_s": len(cfgs) / tg, "batch_gradient_s": tg, "first_call_s": tfirst,
            "forward_batch_s": tfwd, "gradient_over_forward": tg / tfwd, "graph_nodes": nodes,
            "repeat_bitwise_equal": bool(same), "resident_steps_per_box": sorted({r[0] for r in resident}),
            "device_GB": sum(r[2] for r in resident) / 1e9,
            "loss_sum": float(np.sum(J)), "dJ_deps": [float(x) for x in de], "dJ_dq0": [float(x) for x in dq]}


def run_adjoint(steps, ckpt):
    """Gradient of 1/2|psi_T - 0.5 cos(q0 z)|^2 w.r.t. psi0, eps, q0 at 256^3 (checkpoint every 50),
    against the cost of a forward solve of the same length (FP32-mixed)."""
    from smectic3d_amd import Solver
    from smectic3d_amd import _abi as A
    n = 256
    L = n / 8 * 2 * math.pi
    cfg = A.make_config(n, lx=L, ly=L, lz=L, eps=0.3, q0=1.0, beta=0.0, gamma=1.0, dt=0.05,
                        nl_model=A.SMX_NL_CUBIC, precision=A.SMX_FP32_MIXED)
    psi0 = np.random.default_rng(9).uniform(-0.05, 0.05, (n, n, n))
    with Solver(cfg) as s:
        s.autotune(10)
        s.set_l1_tracking(False)
        s.set_psi(psi0)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        tf = time.perf_counter() - t0
        # the forward solve with the same host I/O as the gradient call (psi0 in, one field out)
        t0 = time.perf_counter()
        s.set_psi(psi0)
        s.step(steps)
        s.get_psi()
        tfio = time.perf_counter() - t0
        t0 = time.perf_counter()
        s.adjoint_grad(psi0, steps, ckpt)  # default budget: the workspace is lent for this call only
        tcold = time.perf_counter() - t0
        # the same budget, given explicitly, keeps the workspace in the handle between calls
        s.set_adjoint_budget(s.adjoint_info()[2])
        s.adjoint_grad(psi0, steps, ckpt)
        t0 = time.perf_counter()
        J, d, de, dq = s.adjoint_grad(psi0, steps, ckpt)
        tg = time.perf_counter() - t0
        res, rec, nbytes = s.adjoint_info()
    return {"steps": steps, "ckpt_every": ckpt, "grads_per_s": 1.0 / tg, "forward_s": tf,
            "forward_with_io_s": tfio, "gradient_s": tg,
            "gradient_over_forward": tg / tfio, "gradient_over_forward_steps_only": tg / tf,
            "first_call_s": tcold, "first_call_over_forward": tcold / tfio,
            "resident_steps": res, "recomputed_steps": rec, "device_GB": nbytes / 1e9,
            "note": "gradient_s: a repeated call with an explicit budget (workspace kept in the handle); "
                    "first_call_s: the default budget, which allocates the workspace and releases it on "
                    "return; both include the host transfers of psi0 and dJ/dpsi0, as forward_with_io_s "
                    "includes those of psi0 and psi_T (the forward solve host to host); forward_s times "
                    "the steps alone",
            "loss": J, "dJ_deps": de, "dJ_dq0": dq}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "periodic_256_traffic.json")


def measured_traffic(slot, schedule, n, real="double"):
    """HBM bytes per launch of the kernel behind step slot `slot` under `schedule`, from the
    committed rocprofv3 FETCH_SIZE + WRITE_SIZE summary of an autotuning run (so every schedule
    candidate was dispatched and measured; each counter calibrated against the copy kernel of
    known size); None if that instantiation was not measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            kernels = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    slots = ["k_ypass_inv", "k_zpass_nl", "k_ypass_fwd", "k_xpass_step"]
    if slot not in slots:
        return None
    nc, v = schedule[slots.index(slot)]
    inst = {"k_ypass_inv": f"smx::k_ypass<{real}, {n}, 1, {nc}, {v}>",
            "k_zpass_nl": f"smx::k_zpass<{real}, {n}, {nc}, 0, 1, {v}>",
            "k_ypass_fwd": f"smx::k_ypass<{real}, {n}, -1, {nc}, {v}>",
            "k_xpass_step": f"smx::k_xpass<{real}, {n}, {nc}, 0, {v & 1}, {v >> 1}>"}[slot]
    k = kernels.get(inst)
    if not k or "FETCH_SIZE_bytes" not in k or "WRITE_SIZE_bytes" not in k:
        return None
    return k["FETCH_SIZE_bytes"] + k["WRITE_SIZE_bytes"]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(psi0, warmup=5, steps=20, runs=3):
    """The FP64 oracle (a C port of the reference loop with the reference's dataflow: R2C of psi,
    N, R2C of N, CNAB2, C2R every step; OpenMP threads, its own FFT, not FFTW; one process, no
    MPI) timed on the host on the 256^3 workload: 5 warm-up steps, then 3 runs of 20 timed
    steps, median."""
    from oracle import oracle as O
    from smectic3d_amd import _abi as A
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = O.Oracle(workload_cfg(A.SMX_FP64))
    o.set_psi(psi0)
    o.step(warmup)
    rates = []
    for _ in range(runs):
        t0 = time.perf_counter()
        o.step(steps)
        rates.append(steps / (time.perf_counter() - t0))
    return {"value": float(np.median(rates)), "unit": "steps/s", "cores": threads, "kind": "port",
            "runs_steps_per_s": rates, "cpu_model": cpu_model(), "host_cpus": os.cpu_count(), "processes": 1,
            "fft": "own mixed-radix FP64 (not FFTW)", "mpi": "none",
            "sample": f"256^3 FP64 workload: {warmup} warm-up steps, then {runs} runs of {steps} timed steps, median "
                      f"(oracle/smx_oracle.c, {threads} OpenMP threads, own FFT, not FFTW)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--prof-steps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the accuracy and secondary legs")
    ap.add_argument("--pinned", action="store_true",
                    help="install the schedules of profiles/periodic_256_schedules.json instead of autotuning")
    ap.add_argument("--no-tune", action="store_true", help="default tile schedules (no autotuning)")
    ap.add_argument("--leg", choices=["neumann", "advected", "neumann_advected", "batch", "batch_adjoint",
                                      "adjoint", "accuracy", "surftrack_periodic"],
                    help="run only this secondary workload (with --leg-steps steps) and print its result")
    ap.add_argument("--leg-steps", type=int, default=None)
    ap.add_argument("--schedules-from", default=None,
                    help="install the tile schedules recorded in a bench JSON line or a schedules file")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs the whole domain on a single MI355X "
              "(multi-GPU sharding is a design note, not an implemented path).", file=sys.stderr)
        sys.exit(2)
    from smectic3d_amd import _abi as A

    n = 256
    dx = (n / 8 * 2 * math.pi) / n
    f32 = f64 = None
    sched_src = "autotuned"
    if a.schedules_from:
        prev = json.load(open(a.schedules_from))
        if "fp32" in prev and "fp64" in prev:
            f32, f64 = prev["fp32"], prev["fp64"]
        else:
            f32, f64 = prev["fp32_mixed"]["schedule"], prev["schedule_fp64"]
        sched_src = "from " + a.schedules_from
    elif a.pinned:
        f32, f64 = pinned_schedules()
        if f32 is not None:
            sched_src = "pinned (profiles/periodic_256_schedules.json)"
    if a.no_tune:
        sched_src = "default"

    if a.leg:  # one secondary workload alone (profiling runs)
        runs = {"neumann": (run_neumann, 4000), "advected": (run_advected, 1000),
                "neumann_advected": (run_neumann_advected, 500),
                "batch": (lambda k: run_batch(k, tune=not a.no_tune), 500),
                "batch_adjoint": (lambda k: run_batch_adjoint(k, 50), 500),
                "adjoint": (lambda k: run_adjoint(k, 50), 2000),
                "accuracy": (lambda k: run_accuracy(hyperboloid_ic(n, dx), k, f32, f64), 2000),
                "surftrack_periodic": (run_surftrack_periodic, 20)}
        fn, default_steps = runs[a.leg]
        print(json.dumps({"leg": a.leg, "result": fn(a.leg_steps or default_steps)}, default=str))
        return

    psi0 = hyperboloid_ic(n, dx)
    log(f"workload 256^3, steps={a.steps} warmup={a.warmup}, schedules {sched_src}")
    r64 = run_path(A.SMX_FP64, psi0, a.steps, a.warmup, a.prof_steps, tune=not a.no_tune, fixed=f64)
    log(f"fp64: {r64['steps_per_s']:.1f} steps/s schedule {r64['schedule']}")
    r32 = run_path(A.SMX_FP32_MIXED, psi0, a.steps, a.warmup, a.prof_steps, tune=not a.no_tune, fixed=f32)
    log(f"fp32-mixed: {r32['steps_per_s']:.1f} steps/s schedule {r32['schedule']}")

    def kernel_table(ks):
        return [{"name": k, "avg_ms": m, "alg_bytes": b, "GBs": b / (m * 1e-3) / 1e9} for k, m, b in ks]

    # roofline of the dominant kernel (largest share of the step time) of the FP64 path
    name, ms, alg = max(r64["kernels"], key=lambda k: k[1])
    achieved = alg / (ms * 1e-3) / 1e9
    nzh = n // 2 + 1
    step_alg64 = 2 * n ** 3 * 8 + 5 * n * n * nzh * 16  # compulsory bytes of the pure-SH step DAG
    step_alg32 = step_alg64 // 2
    ksum64 = sum(k[1] for k in r64["kernels"])
    ksum32 = sum(k[1] for k in r32["kernels"])

    line = {
        "metric": METRIC,
        "value": r64["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "periodic SH 256^3, hyperboloid IC + 1e-3 noise, L=32*2pi, q0=1, eps=0.3, beta=0, "
                        "gamma=1, dt=0.1, cubic NL, CNAB2; FP64 (value) with the FP32-mixed path beside it",
            "grid": [n, n, n], "precision": "fp64",
        },
        "schedule_fp64": r64["schedule"],
        "schedule_source": sched_src,
        "schedule_note": "(columns per tile, plan variant) of the y-inverse, z, y-forward and x kernels",
        "kernel_sum_ms_fp64": ksum64,
        "ms_per_step_over_kernel_sum_fp64": r64["ms_per_step"] / ksum64,
        "kernels_fp64": kernel_table(r64["kernels"]),
        "step_roofline_fp64": {"alg_bytes_per_step": step_alg64,
                               "achieved_GBs": step_alg64 / (r64["ms_per_step"] * 1e-3) / 1e9,
                               "frac": step_alg64 / (r64["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "fp32_mixed": {"value": r32["steps_per_s"], "unit": "steps/s", "ms_per_step": r32["ms_per_step"],
                       "schedule": r32["schedule"], "kernel_sum_ms": ksum32,
                       "speedup_vs_fp64": r32["steps_per_s"] / r64["steps_per_s"],
                       "kernels": kernel_table(r32["kernels"]),
                       "step_roofline": {"alg_bytes_per_step": step_alg32,
                                         "achieved_GBs": step_alg32 / (r32["ms_per_step"] * 1e-3) / 1e9,
                                         "frac": step_alg32 / (r32["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS}},
        "copy_kernel_GBs": r32["copy_GBs"],
        "autotune_s": {"fp32": r32["tune_s"], "fp64": r64["tune_s"]},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(name, r64["schedule"], n, "double")},
        "cpu_baseline": None,
    }
    if not a.no_extra:
        log("accuracy: FP32-mixed vs FP64 over 2000 steps ...")
        line["fp32_mixed"]["error_vs_fp64"] = run_accuracy(psi0, 2000, r32["schedule"], r64["schedule"])
        log("batch ...")
        line["batch_24_boxes_fp32"] = run_batch(500, tune=not a.no_tune)
        log("batched gradients ...")
        line["batch_gradients_24_boxes_fp32"] = run_batch_adjoint(500, 50)
        log("neumann + surftrack ...")
        line["neumann_surftrack_256x256x384_fp32"] = run_neumann(4000)
        log("advected ...")
        line["advected_256_fp32"] = run_advected(1000)
        log("advected neumann ...")
        line["advected_neumann_256x256x384_fp32"] = run_neumann_advected(500)
        log("adjoint ...")
        line["adjoint_256_fp32"] = run_adjoint(2000, 50)
    if not a.no_cpu:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(psi0)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
