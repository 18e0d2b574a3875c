"""Benchmark of the periodic Swift-Hohenberg step on one MI355X.

Workload (pf3dMorph-style periodic run): 256^3 grid, L = 32*2*pi per axis
(8 points per wavelength), q0 = 1, eps = 0.3, beta = 0, gamma = 1, dt = 0.1,
cubic nonlinearity, CNAB2; psi0 = 0.4 cos(q0 sqrt(max(x^2+y^2-z^2, 0) + 4))
about the box centre plus uniform noise of amplitude 1e-3 (SplitMix64 seed 0).
Both the FP64 path and the FP32-mixed path are run from the same initial
state; `value` is the FP32-mixed throughput, the FP64 figures and the error of
the mixed path against the FP64 path are reported beside it.

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


def run_path(precision, psi0, steps, warmup, prof_steps, tune=True, fixed=None):
    from smectic3d_amd import Solver
    cfg = workload_cfg(precision)
    s = Solver(cfg)
    t0 = time.perf_counter()
    if fixed is not None:  # schedules of an earlier run (profiled runs use the same kernels)
        s.set_schedule([tuple(x) for x in fixed])
        sched = s.schedule()
    else:
        sched = s.autotune(20) if tune else s.schedule()
    tune_s = time.perf_counter() - t0
    s.set_psi(psi0)
    s.step(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    t1 = time.perf_counter()
    out = {"seconds": t1 - t0, "steps_per_s": steps / (t1 - t0), "ms_per_step": 1e3 * (t1 - t0) / steps,
           "schedule": sched, "tune_s": tune_s}
    out["psi"] = s.get_psi()
    out["energy"] = s.energy()
    out["qdom"] = s.dominant_wavenumber()
    out["l1"] = s.l1()
    out["kernels"] = s.profile_step(prof_steps)  # advances the state; after the final readouts
    if precision == 1:
        # plain float4 copy of one stage array: the achievable-bandwidth yardstick
        n_spec_bytes = cfg.nx * cfg.ny * (cfg.nz // 2 + 1) * 8
        out["copy_GBs"] = 2 * n_spec_bytes / (s.calibration_copy(n_spec_bytes, 50) * 1e-3) / 1e9
    s.close()
    return out


def run_neumann(steps, surf_every=20):
    """cosNoAdvMorph DCT/Neumann + surfTrack: 256x256x384, FP32 bulk with FP64 reductions,
    eps=0.5, q0=1, gamma=1, dt=0.05, hyperboloid IC; surfTrack every 20 steps."""
    from smectic3d_amd import Solver
    from smectic3d_amd import _abi as A
    nx, ny, nz = 256, 256, 384
    dx = 2 * math.pi / 8
    cfg = A.make_config(nx, ny, nz, lx=nx * dx, ly=ny * dx, lz=nz * dx, bc=A.SMX_BC_NEUMANN, eps=0.5, q0=1.0,
                        beta=0.0, gamma=1.0, dt=0.05, nl_model=A.SMX_NL_CUBIC, precision=A.SMX_FP32_MIXED)
    c = (np.arange(nx) - nx // 2) * dx
    cz = (np.arange(nz) - nz // 2) * dx
    r2 = np.maximum(c[:, None, None] ** 2 + c[None, :, None] ** 2 - cz[None, None, :] ** 2, 0.0)
    psi0 = 0.4 * np.cos(np.sqrt(r2 + 4.0))
    with Solver(cfg) as s:
        s.set_psi(psi0)
        s.step(surf_every)
        s.sync()
        t0 = time.perf_counter()
        found = 0
        for _ in range(steps // surf_every):
            s.step(surf_every)
            _, f = s.surftrack(0.7, -1.0)
            found += f
        s.sync()
        t1 = time.perf_counter()
        ks = s.profile_step(20)
    return {"steps": steps, "surftrack_every": surf_every, "steps_per_s_with_surftrack": steps / (t1 - t0),
            "columns_found_per_call": found / max(1, steps // surf_every),
            "kernels": [{"name": k, "avg_ms": m, "GBs": b / (m * 1e-3) / 1e9} for k, m, b in ks]}


def run_advected(steps):
    """Advection-coupled periodic 256^3: mu grad psi force, spectral Stokes (eta=1), eps=0.5, q0=1,
    gamma=1, dt=0.02, psi0 ~ U[-0.05, 0.05]; FP32-mixed."""
    from smectic3d_amd import Solver
    from smectic3d_amd import _abi as A
    n = 256
    L = n / 8 * 2 * math.pi
    cfg = A.make_config(n, lx=L, ly=L, lz=L, eps=0.5, q0=1.0, beta=0.0, gamma=1.0, dt=0.02,
                        nl_model=A.SMX_NL_CUBIC, advect=1, eta=1.0, rho=1.0, force_form=A.SMX_FORCE_MU,
                        precision=A.SMX_FP32_MIXED)
    rng = np.random.default_rng(4)
    with Solver(cfg) as s:
        s.autotune(10)
        s.set_psi(rng.uniform(-0.05, 0.05, (n, n, n)))
        s.step(5)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        t1 = time.perf_counter()
        e = s.energy()
        ks = s.profile_step(10)
    return {"steps": steps, "steps_per_s": steps / (t1 - t0), "energy": e,
            "kernels": [{"name": k, "avg_ms": m, "GBs": b / (m * 1e-3) / 1e9} for k, m, b in ks]}


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
        s.adjoint_grad(psi0, steps, ckpt)  # first call: allocates and first-touches the record sets
        tcold = time.perf_counter() - t0
        t0 = time.perf_counter()
        J, d, de, dq = s.adjoint_grad(psi0, steps, ckpt)
        tg = time.perf_counter() - t0
        res, rec, nbytes = s.adjoint_info()
    return {"steps": steps, "ckpt_every": ckpt, "grads_per_s": 1.0 / tg, "forward_s": tf,
            "forward_with_io_s": tfio, "gradient_s": tg,
            "gradient_over_forward": tg / tfio, "gradient_over_forward_steps_only": tg / tf,
            "first_call_s": tcold, "first_call_over_forward": tcold / tfio,
            "resident_steps": res, "recomputed_steps": rec, "device_GB": nbytes / 1e9,
            "note": "gradient_s: a repeated call (record sets kept in the handle); first_call_s includes "
                    "allocating and first-touching them; both include the host transfers of psi0 and "
                    "dJ/dpsi0, as forward_with_io_s includes those of psi0 and psi_T (the forward solve "
                    "host to host); forward_s times the steps alone",
            "loss": J, "dJ_deps": de, "dJ_dq0": dq}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "periodic_256_fp32_traffic.json")


def measured_traffic(slot, schedule, n):
    """HBM bytes per launch of the kernel behind step slot `slot` under `schedule`, from the
    committed rocprofv3 FETCH_SIZE + WRITE_SIZE summary (each counter calibrated against the
    copy kernel of known size); None if that instantiation was not measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            kernels = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    slots = ["k_ypass_inv", "k_zpass_nl", "k_ypass_fwd", "k_xpass_step"]
    if slot not in slots:
        return None
    nc, v = schedule[slots.index(slot)]
    inst = {"k_ypass_inv": f"smx::k_ypass<float, {n}, 1, {nc}, {v}>",
            "k_zpass_nl": f"smx::k_zpass<float, {n}, {nc}, 0, 1, {v}>",
            "k_ypass_fwd": f"smx::k_ypass<float, {n}, -1, {nc}, {v}>",
            "k_xpass_step": f"smx::k_xpass<float, {n}, {nc}, 0, {v & 1}, {v >> 1}>"}[slot]
    k = kernels.get(inst)
    if not k or "FETCH_SIZE_bytes" not in k or "WRITE_SIZE_bytes" not in k:
        return None
    return k["FETCH_SIZE_bytes"] + k["WRITE_SIZE_bytes"]


def cpu_baseline(psi0, budget_s=12.0):
    """The FP64 oracle (a C port of the reference loop, OpenMP threads) timed on the host,
    on a bounded sample of the same workload: whole steps at 256^3 until ~budget_s."""
    from oracle import oracle as O
    from smectic3d_amd import _abi as A
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = O.Oracle(workload_cfg(A.SMX_FP64))
    o.set_psi(psi0)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.step(1)
        n += 1
        el = time.perf_counter() - t0
        if n >= 2 and el > budget_s:
            break
    return {"value": n / el, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} FP64 steps of the 256^3 workload (oracle/smx_oracle.c, {threads} OpenMP threads, own FFT, not FFTW)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--prof-steps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the Neumann/advected/batch/adjoint legs")
    ap.add_argument("--no-tune", action="store_true", help="default tile schedules (no autotuning)")
    ap.add_argument("--leg", choices=["neumann", "advected", "neumann_advected", "batch", "adjoint"],
                    help="run only this secondary workload (with --leg-steps steps) and print its result")
    ap.add_argument("--leg-steps", type=int, default=None)
    ap.add_argument("--schedules-from", default=None,
                    help="install the tile schedules recorded in an earlier bench JSON line instead of "
                         "autotuning (profiled runs then time the same kernels)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs the whole domain on a single MI355X "
              "(multi-GPU sharding is a design note, not an implemented path).", file=sys.stderr)
        sys.exit(2)
    from smectic3d_amd import _abi as A

    if a.leg:  # one secondary workload alone (profiling runs)
        runs = {"neumann": (run_neumann, 4000), "advected": (run_advected, 1000),
                "neumann_advected": (run_neumann_advected, 500),
                "batch": (lambda k: run_batch(k, tune=not a.no_tune), 500), "adjoint": (lambda k: run_adjoint(k, 50), 2000)}
        fn, default_steps = runs[a.leg]
        print(json.dumps({"leg": a.leg, "result": fn(a.leg_steps or default_steps)}, default=str))
        return

    n = 256
    dx = (n / 8 * 2 * math.pi) / n
    psi0 = hyperboloid_ic(n, dx)
    log(f"workload 256^3, steps={a.steps} warmup={a.warmup}")
    f32 = f64 = None
    if a.schedules_from:
        prev = json.load(open(a.schedules_from))
        f32, f64 = prev["schedule_fp32"], prev["fp64"]["schedule"]
    r32 = run_path(A.SMX_FP32_MIXED, psi0, a.steps, a.warmup, a.prof_steps, tune=not a.no_tune, fixed=f32)
    log(f"fp32-mixed: {r32['steps_per_s']:.1f} steps/s schedule {r32['schedule']}")
    r64 = run_path(A.SMX_FP64, psi0, a.steps, a.warmup, a.prof_steps, tune=not a.no_tune, fixed=f64)
    log(f"fp64: {r64['steps_per_s']:.1f} steps/s")

    err_psi = float(np.linalg.norm(r32["psi"] - r64["psi"]) / np.linalg.norm(r64["psi"]))
    err_e = abs(r32["energy"] - r64["energy"]) / abs(r64["energy"])
    err_q = abs(r32["qdom"] - r64["qdom"]) / abs(r64["qdom"])

    # roofline of the dominant kernel (largest share of the step time) of the FP32-mixed path
    ks = r32["kernels"]
    name, ms, alg = max(ks, key=lambda k: k[1])
    achieved = alg / (ms * 1e-3) / 1e9
    nzh = n // 2 + 1
    s_r, s_c = n ** 3 * 4, n * n * nzh * 8
    step_alg = 2 * s_r + 5 * s_c  # compulsory bytes of the pure-SH step DAG (FP32)

    line = {
        "metric": METRIC,
        "value": r32["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "periodic SH 256^3, hyperboloid IC + 1e-3 noise, L=32*2pi, q0=1, eps=0.3, beta=0, "
                        "gamma=1, dt=0.1, cubic NL, CNAB2 (FP32-mixed headline; FP64 beside it)",
            "grid": [n, n, n], "precision": "fp32-mixed",
        },
        "fp64": {"value": r64["steps_per_s"], "unit": "steps/s", "ms_per_step": r64["ms_per_step"],
                 "schedule": r64["schedule"]},
        "schedule_fp32": r32["schedule"],
        "schedule_note": "(columns per tile, plan variant) of the y-inverse, z, y-forward and x kernels, "
                         "chosen by smx_autotune before the timed region",
        "autotune_s": {"fp32": r32["tune_s"], "fp64": r64["tune_s"]},
        "copy_kernel_GBs": r32["copy_GBs"],
        "speedup_fp32_mixed_vs_fp64": r32["steps_per_s"] / r64["steps_per_s"],
        "error_fp32_mixed_vs_fp64": {"psi_rel_l2": err_psi, "energy_rel": err_e, "qdom_rel": err_q,
                                     "steps": a.steps + a.warmup},
        "step_roofline_fp32": {"alg_bytes_per_step": step_alg,
                               "achieved_GBs": step_alg / (r32["ms_per_step"] * 1e-3) / 1e9,
                               "frac": step_alg / (r32["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "kernels_fp32": [{"name": k, "avg_ms": m, "alg_bytes": b, "GBs": b / (m * 1e-3) / 1e9} for k, m, b in ks],
        "kernels_fp64": [{"name": k, "avg_ms": m, "alg_bytes": b, "GBs": b / (m * 1e-3) / 1e9}
                         for k, m, b in r64["kernels"]],
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(name, r32["schedule"], n)},
        "cpu_baseline": None,
    }
    if not a.no_extra:
        log("batch ...")
        line["batch_24_boxes_fp32"] = run_batch(500, tune=not a.no_tune)
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
