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
    python3 strouhal_amd/profile_workload.py --precision fp32 --copy 50
    python3 strouhal_amd/profile_workload.py --precision fp32 --stream-ceiling
    python3 strouhal_amd/profile_workload.py --precision fp32 --resident

--resident times the multi-lattice batch on lattices one cell deep, with the
streaming kernel (resident mode 0) and with resident launches (mode 2), in
microseconds per step: the ladder's finest rung (1410 x 264: walls, inflow,
NT cylinder, sponge) and the same size with one feature in every cell
(plain periodic, partial cells, sponge, wall links, moving-wall links).

--stream-ceiling times a kernel that moves the step's bytes in the step's
pattern (19 slots read, 19 written, blocked layout, nontemporal) and does no
arithmetic, at 4, 8 and 16 bytes per lane access (libstrouhal_stream.so):
the streaming bound of a step on this lattice.
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


def _stats(ms, first_parity: int = 0) -> dict:
    """Duration statistics of consecutive steps; even / odd: medians of the
    steps taken from an even / odd step count (the two AA passes)."""
    ms = np.asarray(ms, dtype=float)
    ev, od = (ms[0::2], ms[1::2]) if first_parity == 0 else (ms[1::2], ms[0::2])
    return {"steps": int(ms.size), "mean": float(ms.mean()), "median": float(np.median(ms)),
            "min": float(ms.min()), "max": float(ms.max()), "argmax": int(ms.argmax()),
            "even_median": float(np.median(ev)) if ev.size else None,
            "odd_median": float(np.median(od)) if od.size else None,
            "ms": [round(float(x), 4) for x in ms]}


def resident_costs(precision: str, steps: int = 1024) -> dict:
    """us per step of one-lattice batches, streaming (mode 0) vs resident (mode 2)."""
    from strouhal_amd import Simulation
    from strouhal_amd.batch import Batch
    from strouhal_amd.cases import vortex_street
    nx, ny = 1410, 264
    lam_m = 1.5 if precision == "fp32" else None

    def lattice(kind):
        if kind == "street":
            return vortex_street(None, precision, D=64, uin_max=0.5 * 1.5 * 0.1 / math.sqrt(3.0),
                                 lam=3.0 / 16.0 if precision == "fp32" else None).sim
        fl = np.zeros((1, ny, nx), np.uint8)
        if kind == "walls":
            fl[:, :, ::3] = 1       # a solid column every third: every fluid cell has wall links
        if kind == "moving_walls":
            fl[:, :, ::3] = 3       # inflow columns: wall links with velocity corrections
        s = Simulation(None, precision, nx=nx, ny=ny, nz=1, omega=1.9, omega_minus=lam_m, uin=0.05,
                       wants_forces=True, flags=fl)
        s.init_field(1, 0.04)
        if kind == "partial":
            s.set_solid_fractions(np.full((1, ny, nx), 0.3, np.float32))
        if kind == "sponge":
            s.sponge = (-1.0, float(nx), 1.0)
        return s

    out = {}
    for kind in ("street", "plain", "partial", "sponge", "walls", "moving_walls"):
        row = {}
        for mode in (0, 2):
            b = Batch([lattice(kind)])
            b.set_resident(mode)
            b.run([64])
            g0 = b.gpu_seconds
            b.run([steps])
            row["streaming" if mode == 0 else "resident"] = round((b.gpu_seconds - g0) / steps * 1e6, 3)
            b.close()
        out[kind] = row
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16s", "bf16"])
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--dispatch", action="store_true", help="print every step's duration")
    ap.add_argument("--replay", type=int, default=0, help="steps per pass of the snapshot replay")
    ap.add_argument("--copy", type=int, default=0,
                    help="time this many device-to-device copies of the DDF image (a bandwidth ceiling)")
    ap.add_argument("--stream-ceiling", action="store_true",
                    help="time the no-arithmetic 19-slot pass over an n^3 image (see above)")
    ap.add_argument("--resident", action="store_true",
                    help="streaming vs resident batch kernel on lattices one cell deep (see above)")
    ap.add_argument("--case", default="tgv", choices=["tgv", "sphere", "street", "cavity", "channel"],
                    help="tgv: the bench's n^3 Taylor-Green vortex; sphere: the 512x256x256 LES "
                         "sphere (D 40, NT cells, inlets, sponge); street: the ladder's finest "
                         "rung (D 64 at half Mach 0.1); cavity: the 128^2 Re 400 lid-driven "
                         "cavity with its wall frame (SRT); channel: the 512x256 porous channel")
    args = ap.parse_args()
    n = args.n
    if args.resident:
        print(json.dumps({"case": "resident", "precision": args.precision,
                          "us_per_step": resident_costs(args.precision)}), flush=True)
        return
    if args.stream_ceiling:
        import ctypes
        sl = ctypes.CDLL(os.path.join(ROOT, "strouhal_amd", "libstrouhal_stream.so"))
        sl.strh_stream_ceiling.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_float)]
        B = 4 if args.precision == "fp32" else 2
        cells = n ** 3
        out = {"case": "stream_ceiling", "precision": args.precision, "n": n, "bytes_per_pass": 38 * cells * B}
        for lane in (4, 8, 16):
            ms = ctypes.c_float()
            rc = sl.strh_stream_ceiling(cells, B, lane, 50, ctypes.byref(ms))
            if rc != 0:
                raise RuntimeError(f"strh_stream_ceiling failed ({rc})")
            out[f"lane{lane}"] = {"ms": round(ms.value, 4), "TBps": round(38 * cells * B / ms.value / 1e9, 3)}
        print(json.dumps(out), flush=True)
        return
    from strouhal_amd import Simulation, trt_omegas
    lam = 3.0 / 16.0 if args.precision == "fp32" else None
    if args.case == "sphere":
        from strouhal_amd.cases import sphere_case
        sim = sphere_case(None, args.precision, D=40, size=(512, 256, 256), u=0.05, re=300.0,
                          c_smago=0.01, center=(96.0, 128.0, 128.0), lam=None).sim
        n = sim.cells
    elif args.case == "cavity":
        from strouhal_amd.cases import cavity
        sim = cavity(None, args.precision, n=128, re=400.0)
        n = sim.cells
    elif args.case == "channel":
        from strouhal_amd.channel_adjoint import channel_case
        sim, _ = channel_case(None)
        n = sim.cells
    elif args.case == "street":
        from strouhal_amd.cases import vortex_street
        sim = vortex_street(None, args.precision, D=64, uin_max=0.5 * 1.5 * 0.1 / math.sqrt(3.0),
                            lam=lam).sim
        n = sim.cells
    else:
        L = n / (2 * math.pi)
        tau = 3 * (0.05 * L / 1600.0) + 0.5
        wp, wm = trt_omegas(tau, lam)
        sim = Simulation(None, args.precision, nx=n, ny=n, nz=n, omega=wp, omega_minus=wm)
        sim.init_field(2, 0.05)
    head = {"case": args.case, "precision": args.precision, "n": n, "warmup": args.warmup}
    if args.copy > 0:
        # device-to-device copies of the DDF image (strh_snapshot): read and
        # write the same bytes as one step, with no arithmetic and no shift
        import ctypes
        import time
        from strouhal_amd._lib import lib
        sz = ctypes.c_size_t()
        lib().strh_ddf_bytes(sim._h, ctypes.byref(sz))
        sim.snapshot()
        t0 = time.perf_counter()
        for _ in range(args.copy):
            sim.snapshot()
        dt = time.perf_counter() - t0
        print(json.dumps({**head, "copy": {"bytes_per_copy": 2 * sz.value, "copies": args.copy,
                                           "seconds": dt, "GBps": 2 * sz.value * args.copy / dt / 1e9}}),
              flush=True)
        return
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
        par = sim.steps_done & 1
        print(json.dumps({**head, "dispatch": _stats(sim.run_timed(args.steps), par)}), flush=True)
        return
    g0 = sim.gpu_seconds
    sim.run(args.steps + (args.steps & 1))
    gpu_us = (sim.gpu_seconds - g0) / (args.steps + (args.steps & 1)) * 1e6
    mass = sim.mass_sum()
    B = 4 if args.precision == "fp32" else 2
    cells = sim.cells
    print(f"profile_workload: {args.case} {args.precision} {cells} cells, {sim.steps_done} steps, "
          f"{gpu_us:.3f} us per step (HIP events around the run), ddf bytes {19 * cells * B}, "
          f"mass {mass:.6e}", flush=True)


if __name__ == "__main__":
    main()
