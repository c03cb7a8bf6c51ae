"""Shared helpers for the parity tests: seeded states and GPU/oracle pairs."""
from __future__ import annotations

import numpy as np

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17])

STORAGE = {"fp32": "f32", "fp16s": "f16", "bf16": "bf16"}


def random_state(cells: int, seed: int, umax: float = 0.05, noise: float = 2e-4) -> np.ndarray:
    """Shifted DDFs (19, cells) float64: a random equilibrium plus noise."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(-umax, umax, size=(3, cells))
    rhom1 = rng.uniform(-0.01, 0.01, size=cells)
    rho = 1.0 + rhom1
    cu = 3.0 * (CX[:, None] * u[0] + CY[:, None] * u[1] + CZ[:, None] * u[2])
    u2 = (u * u).sum(axis=0)
    feq = W[:, None] * (rhom1 + rho * (cu + 0.5 * cu * cu - 1.5 * u2))
    return feq + rng.normal(0.0, noise, size=feq.shape)


def to_storage(f64: np.ndarray, precision: str) -> np.ndarray:
    """Round a float64 state to the raw storage image of a precision (RNE)."""
    f32 = f64.astype(np.float32)
    if precision == "fp32":
        return f32
    if precision == "fp16s":
        return f32.astype(np.float16)
    u = f32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def raw_bits(a: np.ndarray) -> np.ndarray:
    return a.view(np.uint32) if a.dtype.itemsize == 4 else a.view(np.uint16)


def assert_bit_exact(gpu_img: np.ndarray, ref_img: np.ndarray, what: str = "DDFs"):
    g, r = raw_bits(np.ascontiguousarray(gpu_img)), raw_bits(np.ascontiguousarray(ref_img))
    assert g.shape == r.shape
    bad = np.nonzero(g.reshape(-1) != r.reshape(-1))[0]
    if bad.size:
        k = int(bad[0])
        raise AssertionError(
            f"{bad.size} of {g.size} {what} differ; first at flat index {k} "
            f"(slot {k // g.shape[-1] if g.ndim > 1 else 0}): gpu {gpu_img.reshape(-1)[k]!r} "
            f"oracle {ref_img.reshape(-1)[k]!r}")


def make_pair(gpu, precision, nx, ny, nz, omega, omega_minus=None, flags=None, **kw):
    """A GPU Simulation and an FP32 oracle with identical parameters."""
    from oracle.oracle import OracleSim
    from strouhal_amd import Simulation
    extra = {k: v for k, v in kw.items() if k in ("lid", "uin", "ramp_steps", "force", "c_smago",
                                                  "wants_forces")}
    g = Simulation(gpu, precision, nx=nx, ny=ny, nz=nz, omega=omega, omega_minus=omega_minus,
                   flags=flags, **extra)
    # the oracle receives the float32-rounded values the GPU holds
    r = OracleSim(nx, ny, nz, g.omega, g.omega_minus, arith="f32", storage=STORAGE[precision],
                  lid=g.lid_vel, uin=g.uin_target, ramp_steps=g.ramp_steps, force=g.force,
                  c_smago=g.c_smago, wants_forces=g.wants_forces, flags=g.flags, threads=4)
    return g, r


def dissipation_sum(f: np.ndarray, omega: float, fluid=None) -> float:
    """sum over fluid cells of 2 nu S:S from a natural-order shifted state
    f (19, N), S = -3 Pi_neq / (2 rho tau), nu = (tau - 1/2)/3 (FP64)."""
    f = np.asarray(f, dtype=np.float64)
    tau = 1.0 / omega
    rm1 = f.sum(axis=0)
    rho = 1.0 + rm1
    j = np.stack([(c[:, None] * f).sum(axis=0) for c in (CX, CY, CZ)])
    cc = (CX, CY, CZ)
    ss = np.zeros_like(rho)
    for a in range(3):
        for b in range(3):
            p = (cc[a][:, None] * cc[b][:, None] * f).sum(axis=0) - j[a] * j[b] / rho
            if a == b:
                p = p - rm1 / 3.0
            ss += p * p
    s2 = ss * (1.5 / (rho * tau)) ** 2
    e = 2.0 * (tau - 0.5) / 3.0 * s2
    if fluid is not None:
        e = e[np.asarray(fluid).reshape(-1)]
    return float(e.sum())
