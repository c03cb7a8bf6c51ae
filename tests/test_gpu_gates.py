"""GPU: the reference's gates on the hot path (Sources/strouhal/Cases.swift,
main.swift), restated with the same cases, step counts and thresholds.

* z-periodic cavity: a 3-D lattice with 2-D physics reproduces Ghia Re 400
  and stays z-uniform (Cases.swift:642-686);
* cubic cavity 64^3: mirror symmetry about the mid-z plane (690-725);
* rotation: lid on +x moving +y is the standard cavity transposed (729-765);
* sphere wake demo at Re 100: C_D inside the demo band (1043-1068);
* cavities run to convergence against Ghia (167-279, main.swift:160-193);
* DFG 2D-2 at D = 64 with partial cells, and its digest replay (840-999);
* Taylor-Green Re 1600 against the DNS dissipation peak (1215-1326);
* the credibility gate on the street ladder (1109-1174).

These check the indexing of every axis, the wall and lid links, the
partial-cell forces and the observables at full step counts; parity with
the oracle is the business of test_gpu_parity.py.
"""
from __future__ import annotations

import numpy as np
import pytest

from strouhal_amd import Simulation, trt_omegas
from strouhal_amd.cases import Cell, cavity_flags, sphere_case
from tests.test_oracle_pins import ghia_rms

pytestmark = pytest.mark.gpu

ULID = float(np.float32(0.1))


def _trt_cavity(gpu, n, nz, re, lid, flags, ramp_steps):
    tau = 3.0 * ULID * n / re + 0.5
    wp, wm = trt_omegas(tau, 0.25)
    return Simulation(gpu, "fp32", nx=n + 2, ny=n + 2, nz=nz, omega=wp, omega_minus=wm,
                      lid=lid, ramp_steps=ramp_steps, flags=flags)


def test_cavity_periodic_in_z_is_2d(gpu):
    n, nz = 128, 8
    sim = _trt_cavity(gpu, n, nz, 400.0, (0.1, 0.0, 0.0), cavity_flags(n, nz), 5000)
    sim.run(120_000)
    m = sim.probe_moments().reshape(nz, n + 2, n + 2, 4)
    inner = m[:, 1:n + 1, 1:n + 1, 0]
    z_dev = float(np.abs(inner[1:] - inner[0]).max())
    rms = ghia_rms(m[0].reshape(-1, 4), n, n + 2, ULID, "re400", u_only=True)
    assert np.all(np.isfinite(m))
    assert rms <= 0.02, rms
    assert z_dev <= 1e-5, z_dev


def test_cubic_cavity_mirror_symmetry(gpu):
    n = 64
    N = n + 2
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    fl = np.zeros((N, N, N), dtype=np.uint8)
    wall = (x == 0) | (x == n + 1) | (y == 0) | (y == n + 1) | (z == 0) | (z == n + 1)
    fl[wall] = Cell.solid
    fl[(y == n + 1) & (x >= 1) & (x <= n) & (z >= 1) & (z <= n)] = Cell.lid
    sim = _trt_cavity(gpu, n, N, 400.0, (0.1, 0.0, 0.0), fl, 5000)
    sim.run(120_000)
    m = sim.probe_moments().reshape(N, N, N, 4)
    u = m[1:n + 1, 1:n + 1, 1:n + 1, 0].astype(np.float64)
    asym = float(np.sqrt(np.mean((u - u[::-1]) ** 2))) / ULID
    assert np.all(np.isfinite(m))
    assert asym <= 5e-4, asym


def test_lid_direction_transposed(gpu):
    n = 64
    N = n + 2
    a_fl = cavity_flags(n)                               # lid on the top row, moving +x
    b_fl = np.zeros((1, N, N), dtype=np.uint8)           # lid on the right column, moving +y
    b_fl[0, 0, :] = Cell.solid
    b_fl[0, N - 1, :] = Cell.solid
    b_fl[0, :, 0] = Cell.solid
    b_fl[0, :, N - 1] = Cell.lid
    a = _trt_cavity(gpu, n, 1, 100.0, (0.1, 0.0, 0.0), a_fl, 2000)
    b = _trt_cavity(gpu, n, 1, 100.0, (0.0, 0.1, 0.0), b_fl, 2000)
    a.run(60_000)
    b.run(60_000)
    ma = a.probe_moments().reshape(N, N, 4)[1:n + 1, 1:n + 1].astype(np.float64)
    mb = b.probe_moments().reshape(N, N, 4)[1:n + 1, 1:n + 1].astype(np.float64)
    mbt = mb.transpose(1, 0, 2)                          # cell (x, y) of b at (y, x)
    d2 = (ma[..., 0] - mbt[..., 1]) ** 2 + (ma[..., 1] - mbt[..., 0]) ** 2
    rms = float(np.sqrt(d2.sum() / (2 * n * n))) / ULID
    assert rms <= 5e-4, rms


def test_sphere_wake_demo_re100(gpu):
    c = sphere_case(gpu, "fp32", D=24, size=(192, 96, 96))
    cd, prev = 0.0, np.inf
    while c.sim.steps_done < 40_000:
        c.sim.run(4000 - 2)
        cd = c.drag_coefficient(c.sim.probe_force(c.box_x, c.box_y))
        if c.sim.steps_done > 16_000 and abs(cd - prev) / abs(cd) < 5e-4:
            break
        prev = cd
    assert np.isfinite(cd)
    assert 1.0 <= cd <= 1.45, (cd, c.sim.steps_done)


def _converged_cavity(gpu, precision, n, re, lam, max_steps, check_every=5000, tol=5e-7):
    """The reference's cavity run to convergence (Cases.swift:167-214): every
    check_every steps after the lid ramp, the relative change of the (u, v)
    field; converged below tol, or on a round-off plateau (below 5e-5 in FP32,
    2e-3 otherwise, and no better than 0.95 x the best of the four previous
    checks)."""
    from strouhal_amd.cases import cavity
    sim = cavity(gpu, precision, n=n, re=re, lam=lam)
    prev, history, converged = None, [], False
    while sim.steps_done < max_steps:
        sim.run(min(check_every, max_steps - sim.steps_done))
        cur = sim.probe_moments()[:, :2].astype(np.float64)
        if prev is not None and sim.steps_done > 5000:
            nsum = float(np.sum(cur * cur))
            res = float(np.sqrt(np.sum((cur - prev) ** 2) / nsum)) if nsum > 0 else 0.0
            history.append(res)
            floor = 5e-5 if precision == "fp32" else 2e-3
            if res < tol:
                converged = True
            elif res < floor and len(history) >= 5 and res > 0.95 * min(history[-5:-1]):
                converged = True
        prev = cur
        if converged:
            break
    return sim, converged


@pytest.mark.parametrize("n,re,precision,lam,max_steps", [
    (128, 100.0, "fp32", 0.25, 100_000),        # the smoke gate (main.swift:183-187)
    (256, 1000.0, "fp32", 0.25, 1_200_000),     # the Re 1000 gate (main.swift:160-166, 189-193)
    (256, 1000.0, "fp16s", None, 1_200_000),    # FP16 storage pairs with SRT there
])
def test_cavity_converges_to_ghia(gpu, n, re, precision, lam, max_steps):
    sim, converged = _converged_cavity(gpu, precision, n, re, lam, max_steps)
    m = sim.probe_moments()
    rms = ghia_rms(m, n, n + 2, ULID, {100.0: "re100", 400.0: "re400", 1000.0: "re1000"}[re])
    assert converged, (sim.steps_done, rms)
    assert rms <= 0.02, (rms, sim.steps_done)


def _dfg2(gpu, D, transient, sample_cycles, curved, uin_max=0.075):
    """The reference's DFG 2D-2 run (Cases.swift:840-947): TRT 3/16 channel,
    cylinder at (2D + 1/2, 2D + 1/2) as partial cells (curved) or solid
    cells, C_D / C_L probed every 26 steps after the transient, cycle
    statistics from the upward zero crossings of C_L."""
    from strouhal_amd.cases import channel_flags, vortex_street
    from strouhal_amd.ladder import cycle_stats
    if curved:
        c = vortex_street(gpu, "fp32", D=D, uin_max=uin_max)
        sim, u_mean, box_x, box_y = c.sim, c.u_mean, c.box_x, c.box_y
    else:
        nx, ny = 22 * D + 2, int(4.1 * D) + 2
        u_mean = float(np.float32(uin_max)) * 2.0 / 3.0
        wp, wm = trt_omegas(3.0 * u_mean * D / 100.0 + 0.5, 3.0 / 16.0)
        cx = cy = 0.5 + 2.0 * D
        y, x = np.mgrid[0:ny, 0:nx]
        fl = channel_flags(nx, ny)
        fl[0][((x - cx) ** 2 + (y - cy) ** 2 <= D * D / 4.0) & (y > 0) & (y < ny - 1)] = Cell.solid
        sim = Simulation(gpu, "fp32", nx=nx, ny=ny, nz=1, omega=wp, omega_minus=wm, uin=uin_max,
                         ramp_steps=24_000, wants_forces=True, flags=fl)
        sim.sponge = (float(nx - 1 - 3 * D), float(3 * D), 1.0)
        box_x = (int(cx) - D // 2 - 3, int(cx) + D // 2 + 3)
        box_y = (int(cy) - D // 2 - 3, int(cy) + D // 2 + 3)
    sim.run(transient)
    stride = 26
    samples = int(D / (0.30 * u_mean) * sample_cycles / stride) + 64
    cd, cl, t = [], [], []
    for _ in range(samples):
        sim.run(stride - 2)
        f = sim.probe_force(box_x, box_y)
        cd.append(2.0 * f[0] / (u_mean * u_mean * D))
        cl.append(2.0 * f[1] / (u_mean * u_mean * D))
        t.append(float(sim.steps_done))
    return cycle_stats(cl, cd, t, D, u_mean), sim.state_digest


def test_dfg_2d2_partial_cells_d64(gpu):
    # runDFG2(D: 64, curved: true): transient 60,000 steps, 22 cycles sampled
    st, _ = _dfg2(gpu, 64, 60_000, 22, curved=True)
    assert st["cycles"] >= 15, st
    assert 0.2950 <= st["strouhal"] <= 0.3050, st
    assert 0.9900 <= st["max_cl"] <= 1.0100, st


def test_dfg_2d2_digest_replay(gpu):
    # runDFG2Replay: the unsteady run with its probe schedule, twice
    a = _dfg2(gpu, 40, 30_000, 5, curved=False)
    b = _dfg2(gpu, 40, 30_000, 5, curved=False)
    assert a[1] == b[1]
    assert a[0] == b[0]


def _tgv1600_peak(gpu, n, u0=0.1, t_end=10.5):
    """tgvRun of Cases.swift:1215-1282: TRT Lambda = 1/4 at Re 1600, E(t*)
    every 0.1 t* (per-cell |u|^2 in FP32, summed in FP64 on the GPU),
    eps = -dE/dt* by central differences, parabolic refinement at the peak."""
    u0 = float(np.float32(u0))
    lc = n / (2.0 * np.pi)
    tau = 3.0 * u0 * lc / 1600.0 + 0.5
    wp, wm = trt_omegas(tau, 0.25)
    sim = Simulation(gpu, "fp32", nx=n, ny=n, nz=n, omega=wp, omega_minus=wm)
    sim.init_field(2, u0)
    tc = lc / u0
    dt = 0.1
    stride = max(2, int(tc * dt) & ~1)
    e, t = [], []
    for k in range(int(t_end / dt) + 1):
        if k:
            sim.run(stride)
        e.append(sim.kinetic_energy() / (2.0 * sim.cells * u0 * u0))
        t.append(sim.steps_done / tc)
        assert np.isfinite(e[-1]), (n, t[-1])
    eps = [(t[j], -(e[j + 1] - e[j - 1]) / (t[j + 1] - t[j - 1])) for j in range(1, len(e) - 1)]
    k = max(range(len(eps)), key=lambda j: eps[j][1])
    pt, pe = eps[k]
    if 0 < k < len(eps) - 1:
        y0, y1, y2 = eps[k - 1][1], eps[k][1], eps[k + 1][1]
        den = y0 - 2 * y1 + y2
        if abs(den) > 1e-12:
            d = 0.5 * (y0 - y2) / den
            pt, pe = eps[k][0] + d * dt, y1 - 0.25 * (y0 - y2) * d
    return pt, pe


def test_tgv1600_against_dns(gpu):
    # runTGV1600 (Cases.swift:1301-1326) on its 288^3 / 320^3 pair; the DNS
    # peak eps = 0.012856 at t* = 8.98 (Cases.swift:1214)
    t288, e288 = _tgv1600_peak(gpu, 288)
    t320, e320 = _tgv1600_peak(gpu, 320)
    ref_t, ref_e = 8.98, 0.012856
    assert abs(t320 - ref_t) <= 0.5, (t320, e320)
    assert abs(e320 - e288) / ref_e <= 0.01, (e288, e320)
    assert abs(e320 - ref_e) / ref_e <= 0.10, (t320, e320)


def test_credibility_gate(gpu):
    # runCredibility (Cases.swift:1109-1174): the street ladder on its default
    # rungs D in {32, 48, 64} with the half-Mach anchor at D = 64, full
    # schedule, batched: a finite bar in-domain that covers the published
    # mean C_D (~3.18, featflow level 6) within U + 0.05, every rung recorded.
    import math
    from strouhal_amd.ladder import build, credibility_run, run_batched
    base_umax = 1.5 * 0.1 / math.sqrt(3.0)
    vb = build(gpu, "fp32", [32, 48, 64], 1.0, base_umax)
    series, wall, _, _ = run_batched(vb)
    run = credibility_run(vb, series, wall, base_umax, "fp32")
    b = run.budget
    assert np.isfinite(b.combined) and b.combined > 0, b.headline
    assert abs(b.value - 3.18) <= b.combined + 0.05, b.headline
    assert [r.cells_per_feature for r in run.rungs] == [32, 48, 64]
    assert np.isfinite(run.mach_half)
