"""3-D Taylor-Green vortex at Re 1600: energy and dissipation histories.

The workload of tgvRun (Sources/strouhal/Cases.swift:1217-1283): a triply
periodic n^3 lattice initialised with initField mode 2 at amplitude U0,
nu = U0 L / 1600 on L = n / 2 pi, one convective time t* = L / U0 steps.
Every 0.1 t* (stride max(2, int(0.1 t_c)) & ~1 steps, Cases.swift:1233-1236)
it samples

* E* = sum |u|^2 / (2 N U0^2) (Cases.swift:1242-1243, FP64 on the GPU),
* eps* from -dE*/dt* by central differences (Cases.swift:1254-1257) and
* eps* from the strain rate, 2 nu <S:S> L / U0^3 (FP64 on the GPU),

and reports the dissipation peak (parabolic refinement, Cases.swift:1258-1269)
and the RMS distance to a DNS table over t* in [1, min(t_end, 14)]
(Cases.swift:1271-1278). A run that turns non-finite stops there and says when.

    python3 strouhal_amd/tgv.py --n 256 --precision fp32 --lam 0.1875 --t-end 20
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


def tgv_setup(n: int, u0: float = 0.05, re: float = 1600.0):
    L = n / (2.0 * math.pi)
    nu = u0 * L / re
    tau = 3.0 * nu + 0.5
    tc = L / u0
    stride = max(2, int(tc * 0.1) & ~1)
    return L, tau, tc, stride


def central_eps(times, energy):
    """-dE/dt at interior samples (Cases.swift:1254-1257)."""
    t, e = np.asarray(times), np.asarray(energy)
    return t[1:-1], -(e[2:] - e[:-2]) / (t[2:] - t[:-2])


def peak(times, eps, dt):
    """Largest sample, refined by the parabola through it and its neighbours
    (Cases.swift:1258-1269)."""
    k = int(np.argmax(eps))
    pt, pe = float(times[k]), float(eps[k])
    if 0 < k < len(eps) - 1:
        y0, y1, y2 = eps[k - 1], eps[k], eps[k + 1]
        den = y0 - 2 * y1 + y2
        if abs(den) > 1e-12:
            d = 0.5 * (y0 - y2) / den
            pt = float(times[k] + d * dt)
            pe = float(y1 - 0.25 * (y0 - y2) * d)
    return pt, pe


def rms_vs(times, eps, table_t, table_e, t_end):
    """RMS over table points in [1, t_end] of the nearest sample's distance."""
    times, eps = np.asarray(times), np.asarray(eps)
    s, c = 0.0, 0
    for t, r in zip(table_t, table_e):
        if 1.0 <= t <= t_end and len(times):
            j = int(np.argmin(np.abs(times - t)))
            s += (eps[j] - r) ** 2
            c += 1
    return math.sqrt(s / c) if c else float("nan")


def run(gpu, n: int, precision: str, lam, u0: float = 0.05, t_end: float = 20.0):
    from strouhal_amd import Simulation, trt_omegas
    L, tau, tc, stride = tgv_setup(n, u0)
    wp, wm = trt_omegas(tau, lam)
    sim = Simulation(gpu, precision, nx=n, ny=n, nz=n, omega=wp, omega_minus=wm)
    sim.init_field(2, u0)
    N = n ** 3
    norm_e = 1.0 / (2.0 * N * u0 * u0)
    norm_d = L / (N * u0 ** 3)
    times, energy, eps_s = [], [], []

    def sample():
        times.append(sim.steps_done / tc)
        energy.append(sim.kinetic_energy() * norm_e)
        eps_s.append(sim.dissipation() * norm_d)

    sample()
    samples = int(round(t_end / 0.1))
    nonfinite = None
    t0 = time.perf_counter()
    g0 = sim.gpu_seconds
    for _ in range(samples):
        sim.run(stride)
        sample()
        if not (math.isfinite(energy[-1]) and math.isfinite(eps_s[-1])):
            nonfinite = times[-1]
            break
    wall = time.perf_counter() - t0
    gpu_s = sim.gpu_seconds - g0
    steps = sim.steps_done
    sim.close()
    return {"n": n, "precision": precision, "lambda": lam, "tau": tau, "u0": u0, "stride": stride,
            "steps": steps, "gpu_s": gpu_s, "wall_s": wall,
            "mlups": N * steps / gpu_s / 1e6 if gpu_s > 0 else 0.0,
            "nonfinite_at_t_star": nonfinite, "t": times, "E": energy, "eps_strain": eps_s}


def summarize(r: dict, dns: dict | None):
    t = np.array(r["t"])
    e = np.array(r["E"])
    es = np.array(r["eps_strain"])
    # a decaying vortex loses energy monotonically: the series is cut at the
    # first non-finite sample or the first sample whose energy grows (the
    # onset of an instability), and the statistics use what precedes it
    keep = len(t)
    for k in range(len(t)):
        if not (np.isfinite(e[k]) and np.isfinite(es[k])) or (k and e[k] > e[k - 1]):
            keep = k
            break
    t, e, es = t[:keep], e[:keep], es[:keep]
    out = {k: r[k] for k in ("n", "precision", "lambda", "tau", "steps", "mlups", "nonfinite_at_t_star")}
    out["stable_to_t_star"] = float(t[-1]) if len(t) else 0.0
    out["E0"] = float(e[0])
    out["eps_strain_0"] = float(es[0])
    if len(t) >= 3:
        tc, ec = central_eps(t, e)
        out["peak_dEdt"] = peak(tc, ec, 0.1)
        out["peak_strain"] = peak(t, es, 0.1)
        if dns:
            hi = min(float(t[-1]), 14.0)
            out["rms_dEdt_vs_dns"] = rms_vs(tc, ec, dns["t_star"], dns["eps_star"], hi)
            out["rms_strain_vs_dns"] = rms_vs(t, es, dns["t_star"], dns["eps_star"], hi)
            out["dns_peak"] = [dns["peak"]["t_star"], dns["peak"]["eps_star"]]
    return out


def main():
    import argparse
    from strouhal_amd import default_gpu
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--lam", type=float, default=0.1875, help="TRT Lambda; 0 = SRT")
    ap.add_argument("--u0", type=float, default=0.05)
    ap.add_argument("--t-end", type=float, default=20.0)
    ap.add_argument("--dns", default="", help="JSON table {t_star, eps_star, peak}")
    ap.add_argument("--series", default="", help="write the full series to this JSON")
    args = ap.parse_args()
    lam = None if args.lam == 0 else args.lam
    r = run(default_gpu(), args.n, args.precision, lam, args.u0, args.t_end)
    dns = json.load(open(args.dns)) if args.dns else None
    if args.series:
        json.dump(r, open(args.series, "w"))
    print(json.dumps(summarize(r, dns)), flush=True)


if __name__ == "__main__":
    main()
