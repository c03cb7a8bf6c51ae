#!/usr/bin/env python3
"""MLUPS and HBM-roofline benchmark of the fused D3Q19 collide-and-stream step.

Workload: the triply periodic 3-D Taylor-Green vortex at 256^3 (U0 = 0.05,
Re = 1600 on L = N / 2 pi, tau = 3 nu + 1/2), initialised on the GPU
(initField mode 2, Kernels.metal:414-438) and stepped by the HIP kernels in
two variants: TRT Lambda = 3/16 with FP32 storage (the headline `value`) and
SRT with FP16 shifted storage. A "step" is one LBM time step of all 256^3
cells; the DDFs are resident in HBM before timing starts. The line also
carries the adjoint/forward step-time ratio of the porous-channel objective
(512x256x1, 6000 steps; strh_vjp against strh_run, HIP events), and the
reference's own bench workload and procedure (SRT omega 1.9, 2-D Taylor-Green
init, 20 warm-up and 1,000 timed steps, FP32 and FP16) beside its published
Apple-M5 figures. MLUPS counts every cell, as the reference's bench does
(Sources/strouhal/Cases.swift:141).

The roofline is HBM: algorithmic bytes per cell per step are 19 DDF reads +
19 DDF writes = 152 B (FP32) / 76 B (FP16). The step kernel's average launch
duration comes from HIP events recorded on the library's stream around the
timed steps (the same stream the kernels run on).

cpu_baseline: the FP32 oracle (a scalar C++ restatement of the same update,
oracle/strouhal_oracle.cpp) on the host cores for a bounded sample of the
same workload.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
REFERENCE_M5_FP32_MLUPS = 785.0   # README.md:24-28 of the reference (Apple M5, 256^3 FP32)
REFERENCE_M5_FP16_MLUPS = 1445.0  # README.md:29 of the reference (Apple M5, 256^3 FP16)


def tgv_params(n: int, u0: float = 0.05, re: float = 1600.0):
    L = n / (2.0 * math.pi)
    nu = u0 * L / re
    return 3.0 * nu + 0.5


def measure(precision: str, lam, n: int, steps: int, warmup: int):
    """One leg: `warmup` untimed steps, then exactly `steps` timed steps (wall
    clock around strh_run, which synchronises at its end, and HIP events on the
    library's stream), then the same number of steps again with an event pair
    around every launch for the per-dispatch distribution."""
    from strouhal_amd import Simulation, trt_omegas
    tau = tgv_params(n)
    wp, wm = trt_omegas(tau, lam)
    sim = Simulation(None, precision, nx=n, ny=n, nz=n, omega=wp, omega_minus=wm)
    sim.init_field(2, 0.05)
    sim.run(warmup)
    g0 = sim.gpu_seconds
    t0 = time.perf_counter()
    sim.run(steps)
    wall = time.perf_counter() - t0
    gpu = sim.gpu_seconds - g0
    per = sim.run_timed(max(steps, 2)).astype(float)
    if sim.steps_done & 1:
        sim.run(1)       # the energy probe reads the natural slot order
    ke = sim.kinetic_energy()
    cells = n ** 3
    sim.close()
    return {
        "precision": precision,
        "collision": "SRT" if lam is None else f"TRT(Lambda={lam:g})",
        "tau": tau,
        "mlups": cells * steps / wall / 1e6,
        "mlups_gpu_events": cells * steps / gpu / 1e6,
        "ms_per_step": wall / steps * 1e3,
        "kernel_ms": gpu / steps * 1e3,
        "dispatch_ms": {"steps": int(per.size), "min": float(per.min()), "median": float(np.median(per)),
                        "max": float(per.max()), "mean": float(per.mean()),
                        "first": float(per[0]), "last": float(per[-1])},
        "finite": bool(math.isfinite(ke)),
        "bytes_per_cell": 152 if precision == "fp32" else 76,
    }


def reference_bench(n: int, warmup: int = 20, steps: int = 1000):
    """The reference's own bench workload and procedure (Cases.swift:132-151):
    256^3 periodic, SRT omega = 1.9, initField mode 1
    (2-D Taylor-Green) at A = 0.05, timed with HIP events over `steps` steps
    after `warmup` (the measurement plan asks for >= 20 and >= 1,000), in FP32
    and FP16 storage, beside the reference's published Apple-M5 figures."""
    from strouhal_amd import Simulation
    out = {"workload": f"{n}^3 periodic, SRT omega=1.9, initField mode 1 A=0.05",
           "warmup": warmup, "steps": steps}
    for precision, published in (("fp32", REFERENCE_M5_FP32_MLUPS), ("fp16s", REFERENCE_M5_FP16_MLUPS)):
        sim = Simulation(None, precision, nx=n, ny=n, nz=n, omega=1.9)
        sim.init_field(1, 0.05)
        sim.run(warmup)
        g0 = sim.gpu_seconds
        sim.run(steps)
        gpu = sim.gpu_seconds - g0
        sim.close()
        mlups = n ** 3 * steps / gpu / 1e6
        b = 152 if precision == "fp32" else 76
        out[precision] = {"mlups": round(mlups, 1), "ms_per_step": round(gpu / steps * 1e3, 4),
                          "hbm_frac": round(mlups * 1e6 * b / (HBM_PEAK_GBS * 1e9), 4),
                          "vs_reference_m5": round(mlups / published, 2)}
    return out


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n: int):
    """The oracle (a scalar C++ restatement of the same update) on the host
    cores, in FP32 and FP64 arithmetic, with one thread and with the cores this
    process may use (at most 16, the CPU share of one GPU): a bounded sample of
    the same 256^3 TGV workload (1-2 untimed steps, then 2-6 timed)."""
    from oracle.oracle import OracleSim
    from strouhal_amd import trt_omegas
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    threads = max(1, min(16, avail))
    tau = tgv_params(n)
    wp, wm = trt_omegas(tau, 3.0 / 16.0)
    runs = {}
    for arith in ("f32", "f64"):
        sim = OracleSim(n, n, n, wp, wm, arith=arith, storage=arith, threads=threads)
        sim.init_field(2, 0.05)
        for t, (warm, timed) in ((threads, (2, 6)), (1, (1, 2))):
            sim.set_threads(t)
            sim.run(warm)
            t0 = time.perf_counter()
            sim.run(timed)
            dt = time.perf_counter() - t0
            runs[f"{arith}_{t}t"] = {"mlups": round(n ** 3 * timed / dt / 1e6, 2), "steps": timed,
                                     "seconds": round(dt, 2)}
        del sim
    head = runs[f"f32_{threads}t"]
    return {
        "value": head["mlups"],
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"FP32 oracle, {n}^3 TGV TRT(3/16), {head['steps']} steps after 2 warm-up steps, "
                  f"{threads} OpenMP threads ({head['seconds']} s); FP64 and 1-thread runs in `runs`",
        "runs": runs,
        "nproc": os.cpu_count(),
        "threads_available": avail,
        "cpu_model": _cpu_model(),
    }


def adjoint_ratio(spinup: int = 2000, avg: int = 4000):
    """Adjoint/forward time of the porous-channel objective (512x256x1, Re 50,
    Brinkman design field): strh_vjp against strh_run over the same T steps,
    both from HIP events, with the default checkpoint plan (all moments fit)
    and with ceil(sqrt(T)) checkpoints. ratio_*: the whole VJP (recording
    forward pass, then the reverse pass: segment replays and reverse steps,
    which overlap on two streams when a second recording fits) over the plain
    forward run; sweep_ratio_*: the reverse pass only (what the gradient adds
    to the primal run); reverse_step_ratio_*: the reverse steps alone."""
    from strouhal_amd.channel_adjoint import channel_case
    T = spinup + avg
    obj = dict(inlet_x=8, outlet_x=503, spinup=spinup, avg=avg, w_p=1.0, w_d=1.0)
    fwd, _ = channel_case(None)
    fwd.run(2)
    g0 = fwd.gpu_seconds
    fwd.run(T)
    t_fwd = fwd.gpu_seconds - g0
    out = {"workload": f"porous channel 512x256x1 Re 50, Brinkman design field, T = {T} steps",
           "forward_ms_per_step": round(t_fwd / T * 1e3, 5)}
    for name, k in (("auto", -1), ("sqrt", 0)):
        sim, _ = channel_case(None)
        _, _, _, info = sim.vjp(**obj, checkpoint_every=k)
        t = info["forward_seconds"] + info["reverse_pass_seconds"]
        out[f"ratio_{name}"] = round(t / t_fwd, 3)
        out[f"sweep_ratio_{name}"] = round(info["reverse_pass_seconds"] / t_fwd, 3)
        out[f"reverse_step_ratio_{name}"] = round(info["backward_seconds"] / t_fwd, 3)
        out[f"checkpoint_every_{name}"] = info["checkpoint_every"]
        out[f"workspace_bytes_{name}"] = info["workspace_bytes"]
        sim.close()
    fwd.close()
    return out


def load_traffic():
    p = os.path.join(ROOT, "profiles", "step_kernel_traffic_tgv256.json")
    if not os.path.exists(p):
        return None, None
    with open(p) as fh:
        d = json.load(fh)
    return d.get("fp32_bytes_per_launch"), d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-reference-bench", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU z-slab sharding is a design, not a measured path.", file=sys.stderr)
        sys.exit(2)
    steps, warmup = args.steps, args.warmup
    if steps < 1 or warmup < 0:
        print("bench.py: --steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        sys.exit(2)
    n = args.n

    fp32 = measure("fp32", 3.0 / 16.0, n, steps, warmup)
    fp16 = measure("fp16s", None, n, steps, warmup)

    cells = n ** 3
    achieved = fp32["bytes_per_cell"] * cells / (fp32["kernel_ms"] * 1e-3) / 1e9
    traffic, _ = load_traffic()
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
    }
    fp16_achieved = fp16["bytes_per_cell"] * cells / (fp16["kernel_ms"] * 1e-3) / 1e9
    ref = None if args.no_reference_bench else reference_bench(n)
    adj = None if args.no_adjoint else adjoint_ratio()
    cpu = None if args.no_cpu_baseline else cpu_baseline(n)
    line = {
        "metric": "MLUPS + % HBM roofline, D3Q19 256³ FP32/FP16; adjoint/forward step-time ratio",
        "value": round(fp32["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(fp32["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(fp32["mlups"] / REFERENCE_M5_FP32_MLUPS, 2),
        "dtype": "f32",
        "data": "synthetic (analytic 3-D Taylor-Green initial field)",
        "config": {
            "workload": f"3-D Taylor-Green vortex {n}^3 periodic, U0=0.05, Re=1600, "
                        f"TRT Lambda=3/16, FP32 storage (headline); SRT FP16 shifted storage",
            "lattice": f"D3Q19 {n}x{n}x{n}",
            "cells": cells,
            "tau": round(fp32["tau"], 6),
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp32": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in fp32.items()},
        "fp16s": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in fp16.items()},
        "fp16s_roofline": {"achieved": round(fp16_achieved, 1), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(fp16_achieved / HBM_PEAK_GBS, 4)},
        "adjoint": adj,
        "reference_bench": ref,
        "vs_baseline_ref": "reference README.md:24-28, Apple M5 256^3 FP32 ~785 MLUPS",
        # the reference's own bytes-per-cell model of its bench (19 x 2 DDFs, a
        # 1-byte flag and half of its 8-byte odd-step masks: 157 / 81 B,
        # Cases.swift:142-144), for comparison with its published GB/s
        "reference_byte_model": {
            leg: {"bytes_per_cell": 19 * 2 * b + 1 + 4,
                  "GBps": round(r["mlups"] * (19 * 2 * b + 1 + 4) / 1000.0, 1)}
            for leg, r, b in (("fp32", fp32, 4), ("fp16s", fp16, 2))
        },
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
