"""Short driver for rocprofv3 runs and per-dispatch timings of the step kernel.

Runs the bench workload (3-D Taylor-Green vortex, n^3, TRT 3/16 FP32 or SRT
FP16) for a few steps, bracketed by two kernels of known HBM traffic that
calibrate the FETCH_SIZE / WRITE_SIZE counters:

* init_kernel writes exactly 19 * N * B bytes (and reads nothing);
* the mass reduction's stage-1 kernel reads exactly 19 * N * B bytes.

--dispatch prints the HIP-event duration of every step (strh_run_timed)
after --warmup steps; --replay K additionally snapshots the lattice, runs the
same K steps twice more from the snapshot and then K further steps, printing
each pass's durations (does a slow stretch follow the state or the time since
the load started?).

Usage (under rocprofv3, or alone):
    python3 strouhal_amd/profile_workload.py --precision fp32 --n 256 --steps 10
    python3 strouhal_amd/profile_workload.py --precision fp16s --dispatch --warmup 5 --steps 400
    python3 strouhal_amd/profile_workload.py --precision fp16s --replay 60 --warmup 4
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _stats(ms) -> dict:
    ms = np.asarray(ms, dtype=float)
    return {"steps": int(ms.size), "mean": float(ms.mean()), "median": float(np.median(ms)),
            "min": float(ms.min()), "max": float(ms.max()), "argmax": int(ms.argmax()),
            "ms": [round(float(x), 4) for x in ms]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16s", "bf16"])
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--dispatch", action="store_true", help="print every step's duration")
    ap.add_argument("--replay", type=int, default=0, help="steps per pass of the snapshot replay")
    args = ap.parse_args()
    from strouhal_amd import Simulation, trt_omegas
    n = args.n
    L = n / (2 * math.pi)
    tau = 3 * (0.05 * L / 1600.0) + 0.5
    lam = 3.0 / 16.0 if args.precision == "fp32" else None
    wp, wm = trt_omegas(tau, lam)
    sim = Simulation(None, args.precision, nx=n, ny=n, nz=n, omega=wp, omega_minus=wm)
    sim.init_field(2, 0.05)
    head = {"precision": args.precision, "n": n, "warmup": args.warmup}
    if args.replay > 0:
        sim.run(args.warmup)
        sim.snapshot()
        out = {}
        for rep in ("first", "replay1", "replay2"):
            out[rep] = _stats(sim.run_timed(args.replay))
            sim.restore()
        sim.run(args.replay)
        out["continued"] = _stats(sim.run_timed(args.replay))
        print(json.dumps({**head, "replay": out}), flush=True)
        return
    if args.dispatch:
        sim.run(args.warmup)
        print(json.dumps({**head, "dispatch": _stats(sim.run_timed(args.steps))}), flush=True)
        return
    sim.run(args.steps + (args.steps & 1))
    mass = sim.mass_sum()
    B = 4 if args.precision == "fp32" else 2
    print(f"profile_workload: {args.precision} {n}^3, {sim.steps_done} steps, "
          f"ddf bytes {19 * n ** 3 * B}, mass {mass:.6e}", flush=True)


if __name__ == "__main__":
    main()
