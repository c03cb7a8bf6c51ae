"""Short driver for rocprofv3 runs of the step kernel.

Runs the bench workload (3-D Taylor-Green vortex, n^3, TRT 3/16 FP32 or SRT
FP16) for a few steps, bracketed by two kernels of known HBM traffic that
calibrate the FETCH_SIZE / WRITE_SIZE counters:

* init_kernel writes exactly 19 * N * B bytes (and reads nothing);
* the mass reduction's stage-1 kernel reads exactly 19 * N * B bytes.

Usage (under rocprofv3):
    python3 strouhal_amd/profile_workload.py --precision fp32 --n 256 --steps 10
"""
from __future__ import annotations

import argparse
import math
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16s", "bf16"])
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--steps", type=int, default=10)
    args = ap.parse_args()
    from strouhal_amd import Simulation, trt_omegas
    n = args.n
    L = n / (2 * math.pi)
    tau = 3 * (0.05 * L / 1600.0) + 0.5
    lam = 3.0 / 16.0 if args.precision == "fp32" else None
    wp, wm = trt_omegas(tau, lam)
    sim = Simulation(None, args.precision, nx=n, ny=n, nz=n, omega=wp, omega_minus=wm)
    sim.init_field(2, 0.05)
    sim.run(args.steps + (args.steps & 1))
    mass = sim.mass_sum()
    B = 4 if args.precision == "fp32" else 2
    print(f"profile_workload: {args.precision} {n}^3, {sim.steps_done} steps, "
          f"ddf bytes {19 * n ** 3 * B}, mass {mass:.6e}", flush=True)


if __name__ == "__main__":
    main()
