"""GPU: the reference's 3-D and geometry gates on the hot path, restated with
the same cases, step counts and thresholds (Sources/strouhal/Cases.swift).

* z-periodic cavity: a 3-D lattice with 2-D physics reproduces Ghia Re 400
  and stays z-uniform (Cases.swift:642-686);
* cubic cavity 64^3: mirror symmetry about the mid-z plane (690-725);
* rotation: lid on +x moving +y is the standard cavity transposed (729-765);
* sphere wake demo at Re 100: C_D inside the demo band (1043-1068).

These check the indexing of every axis, the wall and lid links and the
partial-cell forces at full step counts; parity with the oracle is the
business of test_gpu_parity.py.
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
