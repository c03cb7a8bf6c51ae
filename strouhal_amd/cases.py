"""Case geometry and solid-fraction builders (host side).

Restates the inputs the reference builds on the CPU before the hot path runs:

* ``disk_solid_fractions``    Sources/StrouhalCore/CaseSetups.swift:6-25
* ``sphere_solid_fractions``  Sources/StrouhalCore/CaseSetups.swift:68-90
* ``cavity_flags``            Sources/strouhal/Cases.swift:170-180
* ``channel_flags``           Sources/StrouhalCore/CaseSetups.swift:50-54
* ``vortex_street``           Sources/StrouhalCore/CaseSetups.swift:29-65
* ``sphere_case``             Sources/StrouhalCore/CaseSetups.swift:95-134

The sub-sample coordinates are evaluated in the same float64 operation order
as the Swift loops, so the inside-counts (and the fractions) are identical.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .simulation import Cell, Simulation, trt_omegas


def disk_solid_fractions(nx: int, ny: int, cx: float, cy: float, r: float) -> np.ndarray:
    """16x16-supersampled solid fraction of a disk, shape (ny, nx) float32."""
    eps = np.zeros((ny, nx), dtype=np.float32)
    r2 = r * r
    sub = 16
    lo, hi = int(cx - r) - 2, int(cx + r) + 2
    lo_y, hi_y = int(cy - r) - 2, int(cy + r) + 2
    xs = np.arange(max(0, lo), min(nx - 1, hi) + 1, dtype=np.float64)
    ys = np.arange(max(0, lo_y), min(ny - 1, hi_y) + 1, dtype=np.float64)
    if xs.size == 0 or ys.size == 0:
        return eps
    off = (np.arange(sub, dtype=np.float64) + 0.5) / float(sub)
    px = (xs[:, None] - 0.5) + off[None, :]            # (X, sub)
    py = (ys[:, None] - 0.5) + off[None, :]            # (Y, sub)
    dx = px - cx
    dy = py - cy
    inside = (dx[None, :, None, :] * dx[None, :, None, :]
              + dy[:, None, :, None] * dy[:, None, :, None]) <= r2   # (Y, X, sy, sx)
    counts = inside.sum(axis=(2, 3))
    eps[int(ys[0]):int(ys[-1]) + 1, int(xs[0]):int(xs[-1]) + 1] = (
        counts.astype(np.float32) / np.float32(sub * sub))
    return eps


def sphere_solid_fractions(nx: int, ny: int, nz: int, cx: float, cy: float, cz: float,
                           r: float) -> np.ndarray:
    """8x8x8-supersampled solid fraction of a sphere, shape (nz, ny, nx) float32."""
    eps = np.zeros((nz, ny, nx), dtype=np.float32)
    r2 = r * r
    sub = 8
    rng = [np.arange(max(0, int(c - r) - 2), min(n - 1, int(c + r) + 2) + 1, dtype=np.float64)
           for c, n in ((cz, nz), (cy, ny), (cx, nx))]
    zs, ys, xs = rng
    if zs.size == 0 or ys.size == 0 or xs.size == 0:
        return eps
    off = (np.arange(sub, dtype=np.float64) + 0.5) / float(sub)
    dx = ((xs[:, None] - 0.5) + off[None, :]) - cx     # (X, sx)
    dy = ((ys[:, None] - 0.5) + off[None, :]) - cy     # (Y, sy)
    dz = ((zs[:, None] - 0.5) + off[None, :]) - cz     # (Z, sz)
    dx2, dy2, dz2 = dx * dx, dy * dy, dz * dz
    for iz in range(zs.size):
        # (dx^2 + dy^2) + dz^2 in the Swift order
        s = (dx2[None, :, None, None, :] + dy2[:, None, None, :, None])       # (Y, X, 1, sy, sx)
        s = s + dz2[iz][None, None, :, None, None]                               # (Y, X, sz, sy, sx)
        counts = (s <= r2).sum(axis=(2, 3, 4))
        eps[int(zs[iz]), int(ys[0]):int(ys[-1]) + 1, int(xs[0]):int(xs[-1]) + 1] = (
            counts.astype(np.float32) / np.float32(sub ** 3))
    return eps


def cavity_flags(n: int, nz: int = 1) -> np.ndarray:
    """n x n interior + 1-cell solid frame, lid on the top row (Cases.swift:170-180)."""
    nx = ny = n + 2
    fl = np.zeros((nz, ny, nx), dtype=np.uint8)
    fl[:, 0, :] = Cell.solid
    fl[:, :, 0] = Cell.solid
    fl[:, :, nx - 1] = Cell.solid
    fl[:, ny - 1, :] = Cell.lid
    return fl


def channel_flags(nx: int, ny: int, nz: int = 1) -> np.ndarray:
    """Solid walls at y = 0, ny-1; velocity (inflow) walls at x = 0, nx-1."""
    fl = np.zeros((nz, ny, nx), dtype=np.uint8)
    fl[:, :, 0] = Cell.inflow
    fl[:, :, nx - 1] = Cell.inflow
    fl[:, 0, :] = Cell.solid
    fl[:, ny - 1, :] = Cell.solid
    return fl


def cavity(gpu=None, precision="fp32", n: int = 128, re: float = 400.0, lam=None,
           ulid: float = 0.1, ramp_steps: int = 5000) -> Simulation:
    """The lid-driven cavity of Cases.swift:167-180 (nu = u_lid n / Re, tau = 3 nu + 1/2)."""
    nu = float(np.float32(ulid)) * n / re
    tau = 3.0 * nu + 0.5
    wp, wm = trt_omegas(tau, lam)
    return Simulation(gpu, precision, nx=n + 2, ny=n + 2, nz=1, omega=wp, omega_minus=wm,
                      lid=(ulid, 0.0, 0.0), ramp_steps=ramp_steps, flags=cavity_flags(n))


@dataclass
class StreetCase:
    sim: Simulation
    D: int
    u_mean: float
    u_max: float
    box_x: tuple
    box_y: tuple


def vortex_street(gpu=None, precision="fp32", D: int = 40, uin_max: float = 0.075,
                  length_d: int = 22, sponge_d: int = 3, sponge_tau: float = 1.0,
                  re: float = 100.0, wants_forces: bool = True,
                  lam: float | None = 3.0 / 16.0) -> StreetCase:
    """DFG 2D-2 channel with an NT cylinder, sponge and cosine ramp (CaseSetups.swift:37-64).
    lam: TRT magic parameter (the reference's 3/16), None for SRT."""
    nx = length_d * D + 2
    ny = int(4.1 * D) + 2
    u_mean = float(np.float32(uin_max)) * 2.0 / 3.0
    nu = u_mean * D / re
    wp, wm = trt_omegas(3.0 * nu + 0.5, lam)
    cx = 0.5 + 2.0 * D
    cy = 0.5 + 2.0 * D
    sim = Simulation(gpu, precision, nx=nx, ny=ny, nz=1, omega=wp, omega_minus=wm,
                     uin=uin_max, ramp_steps=24_000, wants_forces=wants_forces,
                     flags=channel_flags(nx, ny))
    sim.set_solid_fractions(disk_solid_fractions(nx, ny, cx, cy, D / 2.0))
    sim.sponge = (float(nx - 1 - sponge_d * D), float(sponge_d * D), sponge_tau)
    box_x = (int(cx) - D // 2 - 3, int(cx) + D // 2 + 3)
    box_y = (int(cy) - D // 2 - 3, int(cy) + D // 2 + 3)
    return StreetCase(sim, D, u_mean, uin_max, box_x, box_y)


@dataclass
class SphereCase:
    sim: Simulation
    D: int
    u: float
    re: float
    box_x: tuple
    box_y: tuple
    cd_ref: float

    def drag_coefficient(self, f) -> float:
        """C_D from a force probe, frontal area pi D^2 / 4 (CaseSetups.swift:131-133)."""
        return 2.0 * f[0] / (self.u * self.u * np.pi * self.D * self.D / 4.0)


def sphere_case(gpu=None, precision="fp32", D: int = 48, size=(192, 192, 192),
                u: float = 0.05, re: float = 100.0, c_smago: float = 0.0,
                center=None, lam=3.0 / 16.0) -> SphereCase:
    """Uniform-inflow sphere wake with NT body and outlet sponge (CaseSetups.swift:104-128)."""
    nx, ny, nz = size
    nu = float(np.float32(u)) * D / re
    wp, wm = trt_omegas(3.0 * nu + 0.5, lam)
    if center is None:
        cx, cy, cz = 0.5 + 1.5 * D, ny / 2.0, nz / 2.0
    else:
        cx, cy, cz = center
    fl = np.zeros((nz, ny, nx), dtype=np.uint8)
    fl[:, :, 0] = Cell.inflow
    fl[:, :, nx - 1] = Cell.inflow
    sim = Simulation(gpu, precision, nx=nx, ny=ny, nz=nz, omega=wp, omega_minus=wm, uin=u,
                     ramp_steps=8000, wants_forces=True, c_smago=c_smago, flags=fl)
    sim.inflow_uniform = True
    sim.set_solid_fractions(sphere_solid_fractions(nx, ny, nz, cx, cy, cz, D / 2.0))
    sim.sponge = (float(nx - 1 - D), float(D), 1.0)
    box_x = (int(cx) - D // 2 - 3, int(cx) + D // 2 + 3)
    box_y = (int(cy) - D // 2 - 3, int(cy) + D // 2 + 3)
    cd_ref = 24.0 / re * (1.0 + 0.15 * re ** 0.687)
    return SphereCase(sim, D, float(np.float32(u)), re, box_x, box_y, cd_ref)
