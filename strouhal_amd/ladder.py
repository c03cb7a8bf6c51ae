"""Resolution ladder of the Karman vortex street, sequential and batched.

Restates the schedule and statistics the reference builds around the hot
path:

* ``street_variant``  StreetLadder.variant (Sources/StrouhalCore/CaseSetups.swift:177-201)
* ``probe_schedule``  Ladder.measure's run/probe loop (Sources/StrouhalCore/Ladder.swift:54-71)
* ``cycle_stats``     DFG 2D-2 statistics: Strouhal number from the linearly
                      interpolated upward zero crossings of C_L, max C_D / C_L and
                      mean C_D over complete cycles (Sources/strouhal/Cases.swift:911-947)

``run_sequential`` is the reference's procedure (one lattice after another,
each probe a host round trip); ``run_batched`` advances all lattices in one
batch with the probes on the device. Both produce identical DDFs and
identical force series.

Run as a script for the configuration-3 ladder (D in {16,24,32,48,64} at
Ma 0.1 plus half-Mach anchors) and print one JSON line:

    python3 strouhal_amd/ladder.py --scale 1.0 --precision fp32
"""
from __future__ import annotations

import json
import math
import os
import sys
import time
from dataclasses import dataclass

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from strouhal_amd.batch import Batch  # noqa: E402
from strouhal_amd.cases import vortex_street  # noqa: E402

MA_DEFAULT = 0.1


@dataclass
class Variant:
    D: int
    mach_scale: float
    case: object          # cases.StreetCase
    transient: int
    sample: int
    stride: int

    @property
    def chunk(self) -> int:
        return max(2, self.stride - 2) & ~1

    @property
    def probes(self) -> int:
        per = self.chunk + 2
        return -(-self.sample // per)

    @property
    def total_steps(self) -> int:
        return self.transient + self.probes * (self.chunk + 2)


def street_variant(gpu, precision, D: int, mach_scale: float, base_umax: float,
                   scale_steps: float = 1.0) -> Variant:
    """CaseSetups.swift:177-201 with u_max = base_umax * mach_scale and the
    step schedule scaled by `scale_steps` (1.0 = the reference's). FP32 rungs
    collide with TRT 3/16 (CaseSetups.swift:44-57); 16-bit rungs with SRT, as
    the reference pairs FP16 storage (README.md:78-80: TRT's antisymmetric
    mode amplifies the half-precision quantisation noise)."""
    u = base_umax * mach_scale
    lam = 3.0 / 16.0 if precision == "fp32" else None
    case = vortex_street(gpu, precision, D=D, uin_max=u, re=100.0, lam=lam)
    scale = D / 40.0 / mach_scale
    transient = int(60_000 * scale * scale_steps) & ~1
    sample = int(120_000 * scale * scale_steps) & ~1
    stride = max(8, int(26 * scale)) & ~1
    return Variant(D, mach_scale, case, transient, sample, stride)


def probe_schedule(v: Variant):
    """(first force step, period, count) of Ladder.measure's probes."""
    return v.transient + v.chunk + 1, v.chunk + 2, v.probes


def cycle_stats(cl, cd, t, D: int, u_mean: float) -> dict:
    """Cases.swift:911-947 on sampled C_L / C_D series at steps t."""
    cross = []
    for k in range(1, len(cl)):
        if cl[k - 1] < 0 <= cl[k]:
            frac = -cl[k - 1] / (cl[k] - cl[k - 1])
            cross.append(t[k - 1] + frac * (t[k] - t[k - 1]))
    cycles = max(len(cross) - 1, 0)
    st = 0.0
    max_cd, max_cl, means = 0.0, -math.inf, []
    if cycles >= 2:
        period = (cross[-1] - cross[0]) / cycles
        st = D / (period * u_mean)
        for c in range(cycles):
            lo, hi = cross[c], cross[c + 1]
            sel = [k for k in range(len(cl)) if lo <= t[k] < hi]
            if sel:
                max_cd = max(max_cd, max(cd[k] for k in sel))
                max_cl = max(max_cl, max(cl[k] for k in sel))
                means.append(sum(cd[k] for k in sel) / len(sel))
    return {"strouhal": st, "cycles": cycles, "max_cd": max_cd,
            "max_cl": max_cl if cycles >= 2 else 0.0,
            "mean_cd": float(np.mean(means)) if len(means) >= 1 else 0.0}


def coefficients(v: Variant, forces: np.ndarray):
    um = v.case.u_mean
    cd = 2.0 * forces[:, 0] / (um * um * v.D)
    cl = 2.0 * forces[:, 1] / (um * um * v.D)
    first, period, count = probe_schedule(v)
    t = first + 1 + period * np.arange(len(forces))   # stepsDone after each probe
    return cd, cl, t


def write_qoi_csv(path: str, t, cd, cl):
    """The app's quantity-of-interest export (StrouhalApp.swift:906-915):
    header `step,C_D,C_L`, one `%.0f,%.6f,%.6f` row per sample."""
    with open(path, "w") as fh:
        fh.write("step,C_D,C_L\n")
        for a, b, c in zip(t, cd, cl):
            fh.write("%.0f,%.6f,%.6f\n" % (a, b, c))


def build(gpu, precision, resolutions, scale_steps, base_umax):
    out = []
    for D in resolutions:
        for ms in (1.0, 0.5):
            out.append(street_variant(gpu, precision, D, ms, base_umax, scale_steps))
    return out


def run_sequential(variants):
    """Ladder.measure for each lattice in turn (probe = host round trip)."""
    series = []
    t0 = time.perf_counter()
    for v in variants:
        sim = v.case.sim
        sim.run(v.transient)
        f = []
        for _ in range(v.probes):
            sim.run(v.chunk)
            f.append(sim.probe_force(v.case.box_x, v.case.box_y))
        series.append(np.array(f))
    return series, time.perf_counter() - t0


def run_batched(variants, autotune=True, graph_steps=256, resident=3):
    b = Batch([v.case.sim for v in variants])
    b.set_resident(resident)
    for i, v in enumerate(variants):
        first, period, count = probe_schedule(v)
        b.set_probe(i, first, period, count, v.case.box_x, v.case.box_y)
    bs = b.autotune() if autotune else 256
    b.set_graph_steps(graph_steps)
    t0 = time.perf_counter()
    b.run([v.total_steps for v in variants])
    wall = time.perf_counter() - t0
    series = [b.probe_series(i) for i in range(len(variants))]
    return series, wall, bs, b


def step_rates(gpu, precision, res, base_umax, steps, graph_steps=256, resident=3):
    """Microseconds per lock-step step of every rung together and of the finest
    rung alone, from the rungs' initial state."""
    out = {}
    for sel in ("all", "finest"):
        sims = [v.case.sim for v in build(gpu, precision, res, 0.01, base_umax)]
        if sel == "finest":
            sims = sims[-1:]
        cells = sum(s.cells for s in sims)
        b = Batch(sims)
        b.set_graph_steps(graph_steps)
        b.set_resident(resident)
        b.run([16] * len(sims))
        g0 = b.gpu_seconds
        b.run([steps] * len(sims))
        dt = b.gpu_seconds - g0
        out[sel] = {"us_per_step": dt / steps * 1e6, "glups": cells * steps / dt / 1e9}
        b.close()
    return out


def segment_rates(gpu, precision, res, base_umax, steps, graph_steps=256):
    """Microseconds per lock-step step of each set of rungs that is active
    together in the schedule (every rung, then without those that finish
    first, ...), with the streaming kernel (resident mode 0) and with resident
    launches, time-sliced where needed (mode 2), from the rungs' initial state.
    The schedule's segment lengths weight them into its batched time."""
    vs = build(gpu, precision, res, 0.01, base_umax)
    order = sorted(range(len(vs)), key=lambda i: vs[i].total_steps)
    out = []
    for drop in range(len(vs)):
        if drop > 0 and vs[order[drop - 1]].total_steps == vs[order[drop]].total_steps:
            continue
        act = sorted(order[drop:])
        sims = [vs[i].case.sim for i in act]
        row = {"rungs": [(vs[i].D, vs[i].mach_scale) for i in act]}
        for mode in (0, 2):
            b = Batch(sims)
            b.set_graph_steps(graph_steps)
            b.set_resident(mode)
            b.run([64] * len(sims))
            g0 = b.gpu_seconds
            b.run([steps] * len(sims))
            row["streaming_us" if mode == 0 else "resident_us"] = (b.gpu_seconds - g0) / steps * 1e6
            b.close()
        out.append(row)
    return out


def main():
    import argparse
    from strouhal_amd import default_gpu
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the reference schedule")
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--resolutions", default="16,24,32,48,64")
    ap.add_argument("--mach", type=float, default=MA_DEFAULT)
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--csv", default="", help="directory for one step,C_D,C_L file per rung")
    ap.add_argument("--step-rate", type=int, default=0,
                    help="instead of the schedule: time this many steps of all rungs and of the "
                         "finest rung alone")
    ap.add_argument("--segment-rates", type=int, default=0,
                    help="instead of the schedule: time this many steps of each set of rungs active "
                         "together, streaming and resident")
    ap.add_argument("--graph-steps", type=int, default=256)
    ap.add_argument("--resident", type=int, default=3,
                    help="0: streaming kernel only; 1: resident launches where the active rungs fit on chip "
                         "in one launch; 2: resident launches everywhere (time-sliced groups); 3: as 1, and "
                         "where time-sliced groups would be needed, the faster of the two measured")
    ap.add_argument("--report", default="", help="write the credibility report (markdown) of mean C_D here")
    args = ap.parse_args()
    res = [int(r) for r in args.resolutions.split(",")]
    base_umax = 1.5 * args.mach / math.sqrt(3.0)      # Ma of U_mean = umax * 2/3 * sqrt(3)
    gpu = default_gpu()
    if args.segment_rates > 0:
        print(json.dumps({"workload": f"DFG 2D-2 street ladder D={res}, the rung sets of its schedule, "
                                      f"{args.segment_rates} steps per mode",
                          "precision": args.precision, "graph_steps": args.graph_steps,
                          "segments": segment_rates(gpu, args.precision, res, base_umax, args.segment_rates,
                                                    args.graph_steps)}), flush=True)
        return
    if args.step_rate > 0:
        print(json.dumps({"workload": f"DFG 2D-2 street ladder D={res}, {args.step_rate} steps per mode",
                          "precision": args.precision, "graph_steps": args.graph_steps,
                          "resident": args.resident,
                          **step_rates(gpu, args.precision, res, base_umax, args.step_rate,
                                       args.graph_steps, args.resident)}), flush=True)
        return
    report = {"workload": f"DFG 2D-2 street ladder D={res} x (Ma {args.mach}, half-Mach), "
                          f"schedule x{args.scale}", "precision": args.precision, "resident": args.resident}
    vb = build(gpu, args.precision, res, args.scale, base_umax)
    cells = sum(v.case.sim.cells for v in vb)
    updates = sum(v.case.sim.cells * v.total_steps for v in vb)
    sb, wb, bs, _ = run_batched(vb, resident=args.resident)
    report.update({"lattices": len(vb), "cells": cells, "cell_updates": updates,
                   "batched_s": wb, "batched_mlups": updates / wb / 1e6, "block_size": bs})
    rungs = []
    for v, s in zip(vb, sb):
        cd, cl, t = coefficients(v, s)
        st = cycle_stats(cl, cd, t, v.D, v.case.u_mean)
        if args.csv:
            os.makedirs(args.csv, exist_ok=True)
            write_qoi_csv(os.path.join(args.csv, f"qoi_D{v.D}_ma{v.mach_scale:g}.csv"), t, cd, cl)
        rungs.append({"D": v.D, "mach_scale": v.mach_scale, "steps": v.total_steps,
                      "samples": len(s), "finite": bool(np.all(np.isfinite(s))), **st,
                      "digest": v.case.sim.state_digest[:16]})
    report["rungs"] = rungs
    if not args.no_sequential:
        vs = build(gpu, args.precision, res, args.scale, base_umax)
        ss, ws = run_sequential(vs)
        same = all(np.array_equal(a, b, equal_nan=True) for a, b in zip(sb, ss)) and all(
            a.case.sim.state_digest == b.case.sim.state_digest for a, b in zip(vb, vs))
        report.update({"sequential_s": ws, "speedup": ws / wb, "bitwise_identical": bool(same)})
    if args.report:
        md = credibility_report(vb, sb, rungs, report, wb, base_umax)
        with open(args.report, "w") as fh:
            fh.write(md)
        report["credibility_report"] = args.report
    print(json.dumps(report), flush=True)


def credibility_run(variants, series, wall_s, base_umax, precision):
    """Ladder.credibility (Ladder.swift:74-121) for mean C_D: the Mach-1 rungs
    form the resolution ladder, the finest rung's half-Mach copy is the
    compressibility anchor."""
    from strouhal_amd import credibility as cr
    ladder, half, fine = {}, None, max(v.D for v in variants)
    for v, s in zip(variants, series):
        cd, _, _ = coefficients(v, s)
        if v.mach_scale == 1.0:
            ladder[v.D] = cd
        elif v.D == fine:
            half = cd
    finest = next(v for v in variants if v.D == fine and v.mach_scale == 1.0)
    ma = base_umax * 2.0 / 3.0 * math.sqrt(3.0)           # Ma of the mean inlet speed
    setup = (f"DFG 2D-2 cylinder in a channel 22D x 4.1D, Re 100, parabolic velocity inlet ramped over "
             f"24,000 steps, Noble-Torczynski cylinder, outlet sponge; D in {sorted(ladder)} at Ma {ma:.3f} "
             f"plus a half-Mach anchor at D = {fine}; {precision} storage")
    return cr.credibility("mean C_D", ladder, half, finest.case.sim.state_digest, re=100.0, finest_mach=ma,
                          geometry="bluff body in channel", setup=setup, wall_seconds=wall_s)


def credibility_report(variants, series, rungs, report, wall_s, base_umax) -> str:
    """The credibility report (CredibilityReport.markdown, Ladder.swift:126-202)
    of credibility_run. The verification list holds this run's own checks:
    finite series, the batched / sequential identity when the sequential run
    was made, and the finest rung against the DFG 2D-2 reference intervals."""
    from strouhal_amd import credibility as cr
    fine = max(v.D for v in variants)
    st = next(r for r in rungs if r["D"] == fine and r["mach_scale"] == 1.0)
    gates = [f"every rung's force series finite: {'yes' if all(r['finite'] for r in rungs) else 'NO'}",
             f"finest rung (D = {fine}): St {st['strouhal']:.4f} (DFG 2D-2 interval 0.295-0.305), "
             f"max C_L {st['max_cl']:.4f} (0.99-1.01)"]
    if "bitwise_identical" in report:
        gates.append("batched and sequential ladders bit-identical (every DDF digest, every FP64 force "
                     f"sample): {'yes' if report['bitwise_identical'] else 'NO'}")
    run = credibility_run(variants, series, wall_s, base_umax, report["precision"])
    return cr.markdown(run, gates)

if __name__ == "__main__":
    main()
