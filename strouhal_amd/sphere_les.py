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

With `--adjoint-steps T` each precision instead develops the wake for
`spinup` steps and then differentiates the time-averaged partial-cell drag
(the x momentum exchange of the Noble-Torczynski cells, Kernels.metal:201-227)
over the next T steps with respect to every solid fraction. It times T plain
steps from the same state, the adjoint's recording forward, segment replays
and reverse steps, checks that the adjoint leaves the lattice where the
plain run does, and reports the gradient over the sphere's surface cells
(0 < eps < 1) and the device memory the adjoint holds.

    python3 strouhal_amd/sphere_les.py --spinup 16000 --samples 40 --stride 100
    python3 strouhal_amd/sphere_les.py --spinup 16000 --adjoint-steps 2000 --precisions fp16s,bf16
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


def adjoint_run(gpu, precision: str, size, D: int, spinup: int, T: int, K: int = -1,
                u: float = 0.05, re: float = 300.0, cs: float = 0.1):
    from strouhal_amd._lib import lib
    import ctypes
    nx, ny, nz = size
    case = sphere_case(gpu, precision, D=D, size=size, u=u, re=re, c_smago=cs * cs,
                       center=(96.0, ny / 2.0, nz / 2.0), lam=None)
    sim = case.sim
    sz = ctypes.c_size_t()
    lib().strh_ddf_bytes(sim._h, ctypes.byref(sz))
    ddf = sz.value
    sim.run(spinup)
    sim.snapshot()
    sim.run(2)                                     # warm the kernels at this state
    sim.restore()
    g0 = sim.gpu_seconds
    sim.run(T)
    fwd_s = sim.gpu_seconds - g0
    digest_fwd = sim.state_digest
    sim.restore()
    obj = dict(inlet_x=-1, outlet_x=-1, spinup=0, avg=T, w_p=0.0, w_d=0.0, w_f=1.0)
    t0 = time.perf_counter()
    J, _, gu, info = sim.vjp(**obj, checkpoint_every=K, want_eps=True)
    wall = time.perf_counter() - t0
    same = sim.state_digest == digest_fwd
    eps = np.asarray(sim_eps(case)).reshape(-1)
    ge = info.pop("grad_eps").reshape(-1)
    surf = (eps > 0.0) & (eps < 1.0)
    core = eps >= 1.0
    adj_s = info["forward_seconds"] + info["reverse_pass_seconds"]   # replays and reverse steps may overlap
    n = sim.cells
    # C_D of the averaged partial-cell drag, and its gradient in those units
    cd_scale = 2.0 / (case.u * case.u * np.pi * D * D / 4.0)
    out = {"precision": precision, "spinup": spinup, "adjoint_steps": T, "cells": n,
           "forward_ms_per_step": fwd_s / T * 1e3, "forward_mlups": n * T / fwd_s / 1e6,
           "vjp": info, "vjp_wall_s": wall,
           "adjoint_over_forward": adj_s / fwd_s,
           "recording_forward_ms_per_step": info["forward_seconds"] / T * 1e3,
           "backward_ms_per_step": info["backward_seconds"] / T * 1e3,
           "ddf_bytes": ddf, "adjoint_workspace_bytes": info["workspace_bytes"],
           "device_bytes_during_adjoint": ddf * 2 + info["workspace_bytes"],
           "J_ntF_x": J, "cd_from_J": J * cd_scale, "grad_uin": gu,
           "state_bitwise_equal": bool(same),
           "surface_cells": int(surf.sum()), "core_cells": int(core.sum()),
           "grad_surface_sum": float(ge[surf].sum()), "grad_surface_max_abs": float(np.max(np.abs(ge[surf]))),
           "grad_core_max_abs": float(np.max(np.abs(ge[core]))) if core.any() else 0.0,
           "grad_nonzero_outside_body": int(np.count_nonzero(ge[eps == 0.0])),
           "finite": bool(np.all(np.isfinite(ge)) and np.isfinite(J))}
    sim.close()
    return out


def sim_eps(case):
    nx, ny, nz = case.sim.nx, case.sim.ny, case.sim.nz
    from strouhal_amd.cases import sphere_solid_fractions
    return sphere_solid_fractions(nx, ny, nz, 96.0, ny / 2.0, nz / 2.0, case.D / 2.0)


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
    ap.add_argument("--adjoint-steps", type=int, default=0,
                    help="differentiate the partial-cell drag over this many steps after the spin-up")
    ap.add_argument("--checkpoint-every", type=int, default=-1,
                    help="adjoint segment length; 0 = ceil(sqrt(T)), -1 = the longest that fits")
    args = ap.parse_args()
    gpu = default_gpu()
    size = (args.nx, args.ny, args.nz)
    if args.adjoint_steps > 0:
        for p in args.precisions.split(","):
            r = adjoint_run(gpu, p, size, args.D, args.spinup, args.adjoint_steps, args.checkpoint_every)
            print(json.dumps(r), flush=True)
        return
    res = {}
    for p in args.precisions.split(","):
        res[p] = run(gpu, p, size, args.D, args.spinup, args.samples, args.stride)
        r = res[p]
        print(json.dumps({k: v for k, v in r.items() if k != "cd_series"}), flush=True)
    if "fp32" in res:
        ref = res["fp32"]["cd_mean"]
        for p, r in res.items():
            r["cd_rel_err_vs_fp32"] = abs(r["cd_mean"] - ref) / abs(ref) if ref != 0.0 else None
    report = {"workload": f"sphere D={args.D} in {args.nx}x{args.ny}x{args.nz}, Re 300, U 0.05, "
                          f"SRT + Smagorinsky Cs 0.1, NT solid fractions, sponge; "
                          f"{args.spinup} steps + {args.samples} drag samples every {args.stride}",
              "results": {p: {k: v for k, v in r.items() if k != "cd_series"} for p, r in res.items()}}
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
