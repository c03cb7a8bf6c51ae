"""Differentiable porous channel: forward, objective and adjoint on the GPU.

The workload: a 512x256x1 D3Q19 slab (periodic in z), solid walls at
y = 0 and ny-1, parabolic velocity inlets on both x ends (the channel of
Cases.swift:601-611 with the cylinder replaced by a design field), Re 50 on
the mean inlet velocity and the gap ny-2, TRT with Lambda = 3/16. The
Brinkman porosity is a seeded Gaussian-smoothed random field (sigma = 8
cells) thresholded so that 30% of the cells are porous solid (alpha = 1) and
the rest open (alpha = 0); the first and last `margin` columns stay open.

Objective: the time average over `avg` steps after `spinup` of
w_p (rho_in - rho_out)/3 + w_d sum alpha J_x, with the inlet/outlet density
columns `margin // 2` cells inside each end.

Run as a script to time the plain forward, the objective and the adjoint,
check that the adjoint leaves the lattice bit for bit where the forward
does, and print one JSON line with the gradient at `--sample` seeded cells:

    python3 strouhal_amd/channel_adjoint.py --spinup 2000 --avg 4000
"""
from __future__ import annotations

import json
import math
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from strouhal_amd.cases import channel_flags  # noqa: E402


def _gauss_kernel(sigma: float) -> np.ndarray:
    r = int(math.ceil(4 * sigma))
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum()


def _smooth_axis(a: np.ndarray, k: np.ndarray, axis: int) -> np.ndarray:
    r = len(k) // 2
    pad = [(0, 0)] * a.ndim
    pad[axis] = (r, r)
    p = np.pad(a, pad, mode="reflect")
    return np.apply_along_axis(lambda v: np.convolve(v, k, mode="valid"), axis, p)


def design_field(nx: int, ny: int, sigma: float = 8.0, solid: float = 0.30, seed: int = 0,
                 margin: int = 16) -> np.ndarray:
    """alpha (1, ny, nx) float32 in {0, 1}: the top `solid` fraction of a
    Gaussian-smoothed white-noise field over the interior columns."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((ny, nx))
    k = _gauss_kernel(sigma)
    s = _smooth_axis(_smooth_axis(w, k, 0), k, 1)
    alpha = np.zeros((1, ny, nx), dtype=np.float32)
    inner = s[1:ny - 1, margin:nx - margin]
    thr = np.quantile(inner, 1.0 - solid)
    alpha[0, 1:ny - 1, margin:nx - margin] = (inner > thr).astype(np.float32)
    return alpha


def channel_case(gpu=None, precision: str = "fp32", nx: int = 512, ny: int = 256, re: float = 50.0,
                 umax: float = 0.05, ramp_steps: int = 500, seed: int = 0, margin: int = 16,
                 lam: float = 3.0 / 16.0, nz: int = 1):
    """The channel lattice with its design field (extruded along z when nz > 1)."""
    from strouhal_amd import Simulation, trt_omegas
    u_mean = 2.0 / 3.0 * umax
    nu = u_mean * (ny - 2) / re
    tau = 3.0 * nu + 0.5
    wp, wm = trt_omegas(tau, lam)
    sim = Simulation(gpu, precision, nx=nx, ny=ny, nz=nz, omega=wp, omega_minus=wm, uin=umax,
                     ramp_steps=ramp_steps, flags=channel_flags(nx, ny, nz))
    alpha = np.repeat(design_field(nx, ny, seed=seed, margin=margin), nz, axis=0)
    sim.set_porosity(alpha)
    sim.init_field(0, 0.0)
    return sim, alpha


def sample_cells(alpha: np.ndarray, count: int, seed: int, margin: int):
    """`count` seeded fluid cells (flat indices), half porous and half open."""
    _, ny, nx = alpha.shape
    rng = np.random.default_rng(seed + 1000)
    a = alpha[0]
    ys, xs = np.nonzero(np.ones_like(a, dtype=bool))
    keep = (ys >= 1) & (ys <= ny - 2) & (xs >= margin) & (xs < nx - margin)
    ys, xs = ys[keep], xs[keep]
    por = np.nonzero(a[ys, xs] > 0)[0]
    opn = np.nonzero(a[ys, xs] == 0)[0]
    pick = np.concatenate([rng.choice(por, count // 2, replace=False),
                           rng.choice(opn, count - count // 2, replace=False)])
    return sorted(int(ys[i] * nx + xs[i]) for i in pick)


def main():
    import argparse
    from strouhal_amd import default_gpu
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--nz", type=int, default=1)
    ap.add_argument("--spinup", type=int, default=2000)
    ap.add_argument("--avg", type=int, default=4000)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--checkpoint-every", type=int, default=-1)
    ap.add_argument("--sample", type=int, default=32)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--w-p", type=float, default=1.0)
    ap.add_argument("--w-d", type=float, default=1.0)
    ap.add_argument("--save", default="", help="write the full gradient to this .npy")
    args = ap.parse_args()
    margin = 16
    gpu = default_gpu()
    T = args.spinup + args.avg
    obj = dict(inlet_x=margin // 2, outlet_x=args.nx - 1 - margin // 2, spinup=args.spinup,
               avg=args.avg, w_p=args.w_p, w_d=args.w_d)

    fwd, alpha = channel_case(gpu, args.precision, args.nx, args.ny, seed=args.seed, margin=margin,
                              nz=args.nz)
    fwd.run(2)                                     # warm the kernels
    f0 = fwd.gpu_seconds
    t0 = time.perf_counter()
    fwd.run(T - 2)
    wall_f = time.perf_counter() - t0
    gpu_f = fwd.gpu_seconds - f0

    ev, _ = channel_case(gpu, args.precision, args.nx, args.ny, seed=args.seed, margin=margin,
                              nz=args.nz)
    t0 = time.perf_counter()
    J_eval = ev.objective(**obj)
    wall_o = time.perf_counter() - t0

    adj, _ = channel_case(gpu, args.precision, args.nx, args.ny, seed=args.seed, margin=margin,
                              nz=args.nz)
    t0 = time.perf_counter()
    J, ga, gu, info = adj.vjp(**obj, checkpoint_every=args.checkpoint_every)
    wall_v = time.perf_counter() - t0
    same = adj.state_digest == fwd.state_digest == ev.state_digest

    cells = sample_cells(alpha, args.sample, args.seed, margin)
    g = ga.reshape(-1)
    if args.save:
        np.save(args.save, ga)
    n = adj.cells
    step_ms = gpu_f / (T - 2) * 1e3
    adj_s = info["forward_seconds"] + info["reverse_pass_seconds"]   # replays and reverse steps may overlap
    report = {
        "workload": f"porous channel {args.nx}x{args.ny}x{args.nz} Re 50, Brinkman design field, "
                    f"spin-up {args.spinup} + window {args.avg}",
        "precision": args.precision, "cells": n, "steps": T,
        "forward_ms_per_step": step_ms, "forward_mlups": n / (step_ms * 1e-3) / 1e6,
        "forward_wall_s": wall_f, "objective_wall_s": wall_o, "vjp_wall_s": wall_v,
        "vjp": info, "adjoint_over_forward": adj_s / (gpu_f * T / (T - 2)),
        "adjoint_over_forward_wall": wall_v / (wall_f * T / (T - 2)),
        "backward_ms_per_step": info["backward_seconds"] / T * 1e3,
        "J": J, "J_objective_eval": J_eval, "J_equal": J == J_eval,
        "grad_uin": gu, "state_bitwise_equal": bool(same),
        "grad_norm_inf": float(np.max(np.abs(g))), "grad_sum": float(g.sum()),
        "sample_cells": cells, "sample_grad": [float(g[c]) for c in cells],
        "sample_alpha": [float(alpha.reshape(-1)[c]) for c in cells],
        "seed": args.seed,
    }
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
