"""GPU physics gates at the reference's own sizes (Sources/strouhal/Cases.swift),
plus the config-1 cavity against the FP64 oracle.

Tolerances for the FP32 / FP16 storage runs against the FP64 oracle are
stated per test (CONFIG1_TOL).
"""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

from tests.test_oracle_pins import ghia_rms

pytestmark = pytest.mark.gpu

# Config-1 cavity (130x130x1, SRT, Re 400, u_lid 0.1, 20k steps): maximum
# centre-line velocity difference to the FP64 oracle, in units of u_lid.
CONFIG1_TOL = {"fp32": 2e-4, "fp16s": 2e-2, "bf16": 1e-1}


def _centerlines(m, n, nx):
    u = np.array([(m[iy * nx + n // 2, 0] + m[iy * nx + n // 2 + 1, 0]) / 2.0 for iy in range(1, n + 1)])
    v = np.array([(m[(n // 2) * nx + ix, 1] + m[(n // 2 + 1) * nx + ix, 1]) / 2.0 for ix in range(1, n + 1)])
    return u, v


@pytest.fixture(scope="module")
def config1_oracle():
    from oracle.oracle import OracleSim
    from strouhal_amd import trt_omegas
    from strouhal_amd.cases import cavity_flags
    n, ulid = 128, 0.1
    nu = float(np.float32(ulid)) * n / 400.0
    wp, wm = trt_omegas(3.0 * nu + 0.5, None)
    r = OracleSim(n + 2, n + 2, 1, wp, wm, arith="f64", storage="f64", lid=(ulid, 0, 0),
                  ramp_steps=5000, flags=cavity_flags(n), threads=min(16, os.cpu_count() or 1))
    r.run(20000)
    return r.probe_moments()


@pytest.mark.parametrize("precision", ["fp32", "fp16s", "bf16"])
def test_config1_cavity_vs_fp64_oracle(gpu, config1_oracle, precision):
    """Config 1: lid-driven cavity 128^2 (+frame), SRT, Re=400, 20,000 steps."""
    from strouhal_amd.cases import cavity
    n, ulid = 128, float(np.float32(0.1))
    sim = cavity(gpu, precision, n=n, re=400.0, lam=None)
    sim.run(20000)
    m = sim.probe_moments().astype(np.float64)
    ug, vg = _centerlines(m, n, n + 2)
    ur, vr = _centerlines(config1_oracle, n, n + 2)
    err = max(np.abs(ug - ur).max(), np.abs(vg - vr).max()) / ulid
    print(f"config1 {precision}: max centre-line |u - u_fp64| / u_lid = {err:.3e}")
    assert err <= CONFIG1_TOL[precision]
    d1 = sim.state_digest
    sim2 = cavity(gpu, precision, n=n, re=400.0, lam=None)
    sim2.run(20000)
    assert sim2.state_digest == d1


def test_poiseuille_exact_fp32(gpu):
    """Cases.swift:286-317: TRT 3/16, H=64 (wall-exact scheme). The reference gates the
    FP32 run at 3e-5, its own FP32 rounding floor, which depends on the Metal
    compiler's FMA contraction. Strict-IEEE FP32 (this kernel and the FP32 oracle,
    bit-identical) floors at 7.4e-5; gate 1e-4 here. The FP64 oracle gives ~2e-8."""
    from strouhal_amd import Simulation, trt_omegas
    H, tau, umax = 64, 0.8, 0.05
    nu = (tau - 0.5) / 3.0
    F = 8.0 * nu * umax / (H * H)
    wp, wm = trt_omegas(tau, 3.0 / 16.0)
    ny = H + 2
    fl = np.zeros((1, ny, 16), np.uint8)
    fl[0, 0, :] = fl[0, ny - 1, :] = 1
    sim = Simulation(gpu, "fp32", nx=16, ny=ny, nz=1, omega=wp, omega_minus=wm,
                     force=(F, 0.0, 0.0), flags=fl)
    sim.run((int(6.0 * H * H / nu) + 1) & ~1)
    m = sim.probe_moments()
    err = max(abs(float(m[j * 16 + 8, 0]) - F / (2 * nu) * (j - 0.5) * (H - (j - 0.5))) / umax
              for j in range(1, H + 1))
    print(f"poiseuille H=64 fp32: max err {err:.2e}")
    assert err <= 1e-4


def test_taylor_green_order_fp32(gpu):
    """Cases.swift:329-370: observed order >= 1.9 on N = 32, 64, 128."""
    from strouhal_amd import Simulation, trt_omegas
    nu = 0.02
    wp, wm = trt_omegas(3 * nu + 0.5, 0.25)
    errs = []
    for N in (32, 64, 128):
        u0 = 0.2 * 32.0 / N
        k = 2 * math.pi / N
        steps = int(math.log(2.0) / (2.0 * nu * k * k)) & ~1
        sim = Simulation(gpu, "fp32", nx=N, ny=N, nz=1, omega=wp, omega_minus=wm)
        sim.init_field(1, u0)
        sim.run(steps)
        m = sim.probe_moments().astype(np.float64)
        decay = math.exp(-2.0 * nu * k * k * steps)
        y, x = np.indices((N, N)).reshape(2, -1)
        ue = u0 * decay * np.sin(k * (x + 0.5)) * np.cos(k * (y + 0.5))
        ve = -u0 * decay * np.cos(k * (x + 0.5)) * np.sin(k * (y + 0.5))
        errs.append(math.sqrt(((m[:, 0] - ue) ** 2 + (m[:, 1] - ve) ** 2).sum() / (2 * N * N)) / (u0 * decay))
    orders = [math.log2(errs[i - 1] / errs[i]) for i in (1, 2)]
    print("TG order fp32:", errs, orders)
    assert min(orders) >= 1.9


def test_les_stability_contrast(gpu):
    """Cases.swift:571-605: Re 1e5 cavity 192^2 — bare SRT diverges, TRT + Cs^2=0.04 holds."""
    from strouhal_amd import Simulation, trt_omegas
    from strouhal_amd.cases import cavity_flags

    def max_u(c_smago, lam):
        n, ulid = 192, 0.1
        tau = 3.0 * ulid * n / 1e5 + 0.5
        wp, wm = trt_omegas(tau, lam)
        sim = Simulation(gpu, "fp32", nx=n + 2, ny=n + 2, nz=1, omega=wp, omega_minus=wm,
                         lid=(ulid, 0, 0), ramp_steps=2000, c_smago=c_smago, flags=cavity_flags(n))
        sim.run(30000)
        m = sim.probe_moments()
        if not np.all(np.isfinite(m[:, :2])):
            return float("nan")
        return float(np.abs(m[:, :2]).max())

    bare = max_u(0.0, None)
    les = max_u(0.04, 0.25)
    print(f"LES contrast: bare SRT max|u| {bare}, TRT+LES max|u| {les}")
    assert math.isnan(bare) or bare > 1.0
    assert math.isfinite(les) and les < 0.5


def test_les_laminar_cost(gpu):
    """Cases.swift:610-640: LES on a resolved TG flow stays within max(3x, 5e-3) of LES off."""
    from strouhal_amd import Simulation, trt_omegas
    N, nu, u0 = 64, 0.02, 0.04
    wp, wm = trt_omegas(3 * nu + 0.5, 0.25)
    k = 2 * math.pi / N
    steps = int(math.log(2.0) / (2.0 * nu * k * k)) & ~1

    def rel(cs):
        sim = Simulation(gpu, "fp32", nx=N, ny=N, nz=1, omega=wp, omega_minus=wm, c_smago=cs)
        sim.init_field(1, u0)
        sim.run(steps)
        m = sim.probe_moments().astype(np.float64)
        decay = math.exp(-2.0 * nu * k * k * steps)
        y, x = np.indices((N, N)).reshape(2, -1)
        ue = u0 * decay * np.sin(k * (x + 0.5)) * np.cos(k * (y + 0.5))
        ve = -u0 * decay * np.cos(k * (x + 0.5)) * np.sin(k * (y + 0.5))
        return math.sqrt(((m[:, 0] - ue) ** 2 + (m[:, 1] - ve) ** 2).sum() / (2 * N * N)) / (u0 * decay)

    off, on = rel(0.0), rel(0.01)
    print(f"LES laminar cost: off {off:.2e} on {on:.2e}")
    assert on < max(3 * off, 5e-3)


def test_cavity_mass_drift(gpu):
    """Cases.swift:790-799: |sum(rho-1)| after 50k steps of the Re 100 cavity <= 1e-2."""
    from strouhal_amd.cases import cavity
    sim = cavity(gpu, "fp32", n=128, re=100.0, lam=0.25)
    sim.run(50000)
    drift = abs(sim.mass_sum())
    print(f"lid mass drift: {drift:.2e}")
    assert drift <= 1e-2


def test_dfg_2d1_nt_drag(gpu):
    """Cases.swift:446-533 (NT curved, D=40): |C_D - 5.5795| / 5.5795 <= 1%."""
    from strouhal_amd import Simulation, trt_omegas
    from strouhal_amd.cases import channel_flags, disk_solid_fractions
    D, uin = 40, 0.075
    nx, ny = int(22.0 * D) + 2, int(4.1 * D) + 2
    u_mean = float(np.float32(uin)) * 2.0 / 3.0
    nu = u_mean * D / 20.0
    wp, wm = trt_omegas(3 * nu + 0.5, 3.0 / 16.0)
    cx = cy = 0.5 + 2.0 * D
    sim = Simulation(gpu, "fp32", nx=nx, ny=ny, nz=1, omega=wp, omega_minus=wm, uin=uin,
                     ramp_steps=4000, wants_forces=True, flags=channel_flags(nx, ny))
    sim.set_solid_fractions(disk_solid_fractions(nx, ny, cx, cy, D / 2.0))
    bx = (int(cx) - D // 2 - 3, int(cx) + D // 2 + 3)
    by = (int(cy) - D // 2 - 3, int(cy) + D // 2 + 3)
    flow_through = int(nx / u_mean)
    prev, settled, cd = float("inf"), 0, 0.0
    while sim.steps_done < 240_000:
        sim.run(4000 - 2)
        f = sim.probe_force(bx, by)
        cd = 2.0 * f[0] / (u_mean * u_mean * D)
        if sim.steps_done > 3 * flow_through:
            settled = settled + 1 if abs(cd - prev) / abs(cd) < 1e-5 else 0
            if settled >= 3:
                break
        prev = cd
    err = abs(cd - 5.57953523384) / 5.57953523384
    print(f"DFG 2D-1 NT D=40: C_D {cd:.4f} (err {err * 100:.2f}%) after {sim.steps_done} steps")
    assert err <= 0.01


def test_tgv_256_energy_decay(gpu):
    """Config 2 (short): 256^3 TGV, U0=0.05, Re=1600, TRT 3/16 FP32 and SRT FP16,
    2 convective times; E(0) matches the analytic 1/8 within the initial-field
    discretisation error, E decays monotonically, and the FP16 run tracks FP32."""
    from strouhal_amd import Simulation, trt_omegas
    n = 256
    L = n / (2 * math.pi)
    nu = 0.05 * L / 1600.0
    tau = 3 * nu + 0.5
    tc = L / 0.05
    stride = int(tc * 0.1) & ~1
    series = {}
    for precision, lam in (("fp32", 3.0 / 16.0), ("fp16s", None)):
        wp, wm = trt_omegas(tau, lam)
        sim = Simulation(gpu, precision, nx=n, ny=n, nz=n, omega=wp, omega_minus=wm)
        sim.init_field(2, 0.05)
        e = [sim.kinetic_energy() / (2.0 * n ** 3 * 0.05 ** 2)]
        for _ in range(20):
            sim.run(stride)
            e.append(sim.kinetic_energy() / (2.0 * n ** 3 * 0.05 ** 2))
        series[precision] = np.array(e)
        sim.close()
    e32, e16 = series["fp32"], series["fp16s"]
    print("TGV E(t*) fp32:", np.round(e32[::5], 6), "fp16:", np.round(e16[::5], 6))
    assert abs(e32[0] - 0.125) < 1e-3
    assert np.all(np.diff(e32) < 0)
    assert np.max(np.abs(e16 - e32)) / e32[0] < 5e-3


def test_tgv_512_energy_decay(gpu):
    """Config 2 at its full 512^3 size: U0=0.05, Re=1600 (tau = 0.507639), TRT
    3/16 with FP32 storage (10.2 GB of DDFs, the blocked image beyond 4 GiB)
    and SRT with FP16 shifted storage, 0.5 convective times sampled every
    0.1 t*: E*(0) = 1/8 within the initial-field discretisation error, E*
    decays monotonically, the FP16 run tracks FP32 within 5e-3 of E*(0), and
    the mean decay rate over 0.5 t* is within 5% of the initial field's
    viscous dissipation, dE*/dt* = -2 (3 / Re) E*(0) = -3 / (4 Re) (k^2 = 3 in
    units of L; the laminar stage of the vortex)."""
    from strouhal_amd import Simulation, trt_omegas
    from strouhal_amd.tgv import tgv_setup
    n, u0 = 512, 0.05
    L, tau, tc, stride = tgv_setup(n, u0)
    assert abs(tau - 0.507639) < 1e-6
    series = {}
    for precision, lam in (("fp32", 3.0 / 16.0), ("fp16s", None)):
        wp, wm = trt_omegas(tau, lam)
        sim = Simulation(gpu, precision, nx=n, ny=n, nz=n, omega=wp, omega_minus=wm)
        sim.init_field(2, u0)
        e = [sim.kinetic_energy() / (2.0 * n ** 3 * u0 ** 2)]
        for _ in range(5):
            sim.run(stride)
            e.append(sim.kinetic_energy() / (2.0 * n ** 3 * u0 ** 2))
        assert sim.steps_done == 5 * stride
        series[precision] = np.array(e)
        sim.close()
    e32, e16 = series["fp32"], series["fp16s"]
    print("TGV 512^3 E*(t*) fp32:", np.round(e32, 7), "fp16:", np.round(e16, 7))
    assert np.all(np.isfinite(e32)) and np.all(np.isfinite(e16))
    assert abs(e32[0] - 0.125) < 1e-3
    assert np.all(np.diff(e32) < 0) and np.all(np.diff(e16) < 0)
    assert np.max(np.abs(e16 - e32)) / e32[0] < 5e-3
    # early laminar decay of the TGV: dE*/dt* = -2 (3/Re) E*(0) at t* -> 0
    t = 5 * stride / tc
    rate = (e32[0] - e32[-1]) / t
    assert rate == pytest.approx(2.0 * 3.0 / 1600.0 * e32[0], rel=0.05), rate


@pytest.mark.parametrize("precision,lam", [("fp32", 3.0 / 16.0), ("fp16s", None), ("bf16", None)])
def test_tgv64_energy_and_dissipation_vs_fp64_oracle(gpu, precision, lam):
    """Config 2's observables at 64^3 over one convective time, every 0.1 t*:
    E* and the strain-rate eps* (both FP64 GPU reductions) against the FP64
    oracle, with the error per precision stated in the assertion bounds."""
    from oracle.oracle import OracleSim
    from strouhal_amd import Simulation, trt_omegas
    from strouhal_amd.tgv import tgv_setup
    from tests.helpers import dissipation_sum
    n, u0 = 64, 0.05
    L, tau, tc, stride = tgv_setup(n, u0)
    wp, wm = trt_omegas(tau, lam)
    g = Simulation(gpu, precision, nx=n, ny=n, nz=n, omega=wp, omega_minus=wm)
    r = OracleSim(n, n, n, g.omega, g.omega_minus, arith="f64", storage="f64", threads=8)
    g.init_field(2, u0)
    r.init_field(2, u0)
    N = n ** 3
    ne, nd = 1.0 / (2 * N * u0 * u0), L / (N * u0 ** 3)
    err_e, err_d = [], []
    for k in range(11):
        if k:
            g.run(stride)
            r.run(stride)
        m = r.probe_moments()
        e_ref = float((m[:, :3] ** 2).sum()) * ne
        d_ref = dissipation_sum(r.download_ddf().reshape(19, -1), g.omega) * nd
        err_e.append(abs(g.kinetic_energy() * ne - e_ref) / e_ref)
        err_d.append(abs(g.dissipation() * nd - d_ref) / d_ref)
    print(f"TGV 64^3 {precision}: max rel err E {max(err_e):.2e}, eps {max(err_d):.2e}")
    # measured on MI355X: fp32 1e-7 / 1.3e-7, fp16s 8e-6 / 6e-4, bf16 1.2e-3 / 3.4e-2
    bound = {"fp32": (1e-6, 1e-5), "fp16s": (1e-4, 5e-3), "bf16": (1e-2, 0.1)}[precision]
    assert max(err_e) < bound[0] and max(err_d) < bound[1]
