"""3-D sphere wake with Smagorinsky LES in FP32, FP16 and BF16 storage.

The workload: a 512x256x256 lattice, a sphere of diameter 40 cells centred
at (96, 128, 128) as Noble-Torczynski solid fractions (CaseSetups.swift:68-90),
uniform velocity inlets on both x ends, U = 0.05, Re = 300, SRT with
Smagorinsky Cs = 0.1 (cSmago = Cs^2 = 0.01, Kernels.metal:54), an outlet
sponge over the last D cells and the inlet ramped over 8000 steps (the
sphere case of CaseSetups.swift:104-128 at this size).

Each precision runs `spinup` steps, then samples the momentum-exchange drag
every `stride` steps `samples` times (FP64 box sums on the GPU,
Sim.swift:359-372) and reports the mean C_D over the samples, the Schiller-
Naumann correlation, MLUPS and the device memory the lattice holds; FP16 and
BF16 are compared with FP32.

    python3 strouhal_amd/sphere_les.py --spinup 16000 --samples 40 --stride 100
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from strouhal_amd.cases import sphere_case  # noqa: E402


def run(gpu, precision: str, size, D: int, spinup: int, samples: int, stride: int,
        u: float = 0.05, re: float = 300.0, cs: float = 0.1):
    nx, ny, nz = size
    case = sphere_case(gpu, precision, D=D, size=size, u=u, re=re, c_smago=cs * cs,
                       center=(96.0, ny / 2.0, nz / 2.0), lam=None)
    sim = case.sim
    nbytes = 0
    from strouhal_amd._lib import lib
    import ctypes
    sz = ctypes.c_size_t()
    lib().strh_ddf_bytes(sim._h, ctypes.byref(sz))
    nbytes = sz.value
    g0 = sim.gpu_seconds
    t0 = time.perf_counter()
    sim.run(spinup)
    cd = []
    for _ in range(samples):
        sim.run(stride - 2)                 # the probe advances 2 steps itself
        cd.append(case.drag_coefficient(sim.probe_force(case.box_x, case.box_y)))
    wall = time.perf_counter() - t0
    gpu_s = sim.gpu_seconds - g0
    steps = sim.steps_done
    out = {"precision": precision, "steps": steps, "gpu_s": gpu_s, "wall_s": wall,
           "mlups": sim.cells * steps / gpu_s / 1e6,
           "ddf_bytes": nbytes, "cd_mean": float(np.mean(cd)), "cd_std": float(np.std(cd)),
           "cd_series": [float(c) for c in cd], "cd_schiller_naumann": case.cd_ref,
           "finite": bool(np.all(np.isfinite(cd)))}
    sim.close()
    return out


def main():
    import argparse
    from strouhal_amd import default_gpu
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--D", type=int, default=40)
    ap.add_argument("--spinup", type=int, default=16000)
    ap.add_argument("--samples", type=int, default=40)
    ap.add_argument("--stride", type=int, default=100)
    ap.add_argument("--precisions", default="fp32,fp16s,bf16")
    args = ap.parse_args()
    gpu = default_gpu()
    size = (args.nx, args.ny, args.nz)
    res = {}
    for p in args.precisions.split(","):
        res[p] = run(gpu, p, size, args.D, args.spinup, args.samples, args.stride)
        r = res[p]
        print(json.dumps({k: v for k, v in r.items() if k != "cd_series"}), flush=True)
    if "fp32" in res:
        ref = res["fp32"]["cd_mean"]
        for p, r in res.items():
            r["cd_rel_err_vs_fp32"] = abs(r["cd_mean"] - ref) / abs(ref)
    report = {"workload": f"sphere D={args.D} in {args.nx}x{args.ny}x{args.nz}, Re 300, U 0.05, "
                          f"SRT + Smagorinsky Cs 0.1, NT solid fractions, sponge; "
                          f"{args.spinup} steps + {args.samples} drag samples every {args.stride}",
              "results": {p: {k: v for k, v in r.items() if k != "cd_series"} for p, r in res.items()}}
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
