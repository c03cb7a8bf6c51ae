"""Full-size check of the porous-channel adjoint (run by hand; not collected).

Reads the JSON line `strouhal_amd/channel_adjoint.py` printed on the GPU (and,
if given, the full gradient it saved with --save), rebuilds the same lattice
in the FP64 oracle and computes

* the FP64 discrete adjoint (checkpointed, all host threads),
* central finite differences of the FP64 objective at the report's sample
  cells (h = 1e-4 on alpha, measured after float32 rounding) and for the
  inlet velocity (h = 1e-6),

then prints one JSON line comparing them with the GPU values. At 512x256 and
6000 steps one objective evaluation takes about 45 s on 8 cores, so the
whole check takes about an hour.

    python3 tests/fullsize_channel_adjoint.py gpu_report.json [gradient.npy] [--threads 8]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle.oracle import OracleSim  # noqa: E402
from strouhal_amd.cases import channel_flags  # noqa: E402
from strouhal_amd.channel_adjoint import design_field  # noqa: E402
from strouhal_amd.simulation import trt_omegas  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("report")
    ap.add_argument("gradient", nargs="?", default="")
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--spinup", type=int, default=2000)
    ap.add_argument("--avg", type=int, default=4000)
    ap.add_argument("--skip-fd", action="store_true")
    args = ap.parse_args()
    rep = json.loads(open(args.report).read().strip().splitlines()[-1])
    nx, ny, margin = args.nx, args.ny, 16
    umax, re, ramp = 0.05, 50.0, 500
    tau = 3.0 * (2.0 / 3.0 * umax * (ny - 2) / re) + 0.5
    wp, wm = trt_omegas(tau, 3.0 / 16.0)
    alpha = design_field(nx, ny, seed=rep["seed"], margin=margin)
    obj = dict(inlet_x=margin // 2, outlet_x=nx - 1 - margin // 2, spinup=args.spinup, avg=args.avg,
               w_p=1.0, w_d=1.0)
    u32 = float(np.float32(umax))

    def make(al=alpha, u=u32):
        s = OracleSim(nx, ny, 1, wp, wm, arith="f64", storage="f64", uin=u, ramp_steps=ramp,
                      flags=channel_flags(nx, ny, 1), threads=args.threads)
        s.set_porosity(al)
        s.init_field(0, 0.0)
        return s

    out = {"workload": rep["workload"], "threads": args.threads}
    t0 = time.time()
    J64, g64, gu64 = make().vjp(**obj)
    out["oracle_adjoint_s"] = time.time() - t0
    out["J_gpu"], out["J_fp64"] = rep["J"], J64
    out["J_rel_err"] = abs(rep["J"] - J64) / abs(J64)
    out["grad_uin_gpu"], out["grad_uin_fp64_adjoint"] = rep["grad_uin"], gu64
    cells = rep["sample_cells"]
    gs = np.array(rep["sample_grad"])
    ga = g64[cells]
    out["sample_fp64_adjoint"] = ga.tolist()
    out["sample_err_vs_fp64_adjoint_over_max"] = float(np.max(np.abs(gs - ga)) / np.max(np.abs(ga)))
    if args.gradient:
        full = np.load(args.gradient).reshape(-1)
        out["field_err_vs_fp64_adjoint_over_max"] = float(np.max(np.abs(full - g64)) / np.max(np.abs(g64)))
    print(json.dumps(out), flush=True)
    if args.skip_fd:
        return
    h = 1e-4
    fd = []
    for c in cells:
        ap_, am_ = alpha.copy().reshape(-1), alpha.copy().reshape(-1)
        ap_[c] = np.float32(alpha.reshape(-1)[c] + h)
        am_[c] = np.float32(alpha.reshape(-1)[c] - h)
        d = float(ap_[c]) - float(am_[c])
        Jp = make(ap_.reshape(alpha.shape)).objective(**obj)
        Jm = make(am_.reshape(alpha.shape)).objective(**obj)
        fd.append((Jp - Jm) / d)
        print(json.dumps({"cell": c, "fd": fd[-1], "gpu": float(gs[len(fd) - 1]),
                          "fp64_adjoint": float(ga[len(fd) - 1])}), flush=True)
    fd = np.array(fd)
    hu = 1e-6
    fdu = (make(u=u32 + hu).objective(**obj) - make(u=u32 - hu).objective(**obj)) / (2 * hu)
    big = np.abs(fd) >= 1e-3 * np.max(np.abs(fd))
    out.update({
        "fd_h": h, "sample_fd": fd.tolist(), "grad_uin_fd": fdu,
        "gpu_vs_fd_err_over_max": float(np.max(np.abs(gs - fd)) / np.max(np.abs(fd))),
        "fp64_adjoint_vs_fd_err_over_max": float(np.max(np.abs(ga - fd)) / np.max(np.abs(fd))),
        "gpu_vs_fd_max_rel_err_large_cells": float(np.max(np.abs(gs[big] - fd[big]) / np.abs(fd[big]))),
        "large_cells": int(big.sum()),
        "grad_uin_gpu_vs_fd_rel_err": abs(rep["grad_uin"] - fdu) / abs(fdu),
        "total_s": time.time() - t0,
    })
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
