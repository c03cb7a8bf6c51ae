"""Host-side mirror of StrouhalCore's solver API over the HIP C ABI.

Names, argument meaning and error behaviour follow the reference's Swift
surface (Sources/StrouhalCore/Sim.swift):

* ``Precision``  Sim.swift:5-8  (``fp32``, ``fp16s``; ``bf16`` is new)
* ``Cell``       Sim.swift:46-52
* ``GPU``        Sim.swift:60-125 (one HIP context + stream per instance)
* ``Simulation`` Sim.swift:153-459
* ``trt_omegas`` Sim.swift:188-193

Differences forced by discrete HBM: ``probe_moments`` returns a host copy
(the reference returns a view of shared memory, Sim.swift:351-352), and the
odd-step preconditions (Sim.swift:335, 360) raise ``StrouhalError`` with
``STRH_ERR_ODD_STEPS`` instead of trapping.
"""
from __future__ import annotations

import ctypes
import os
import enum

import numpy as np

from ._lib import LatticeDesc, StrouhalError, check, lib


class Precision(str, enum.Enum):
    fp32 = "fp32"
    fp16s = "fp16s"
    bf16 = "bf16"

    @property
    def ddf_bytes(self) -> int:
        return 4 if self is Precision.fp32 else 2

    @property
    def code(self) -> int:
        return {"fp32": 0, "fp16s": 1, "bf16": 2}[self.value]

    @property
    def dtype(self):
        return np.float32 if self is Precision.fp32 else (np.float16 if self is Precision.fp16s else np.uint16)


class Cell(enum.IntEnum):
    fluid = 0
    solid = 1
    lid = 2       # moving wall, velocity lidVel
    inflow = 3    # moving wall, parabolic (or uniform) +x profile
    outflow = 4   # dormant anti-bounce-back wall (Kernels.metal:28-29)


def trt_omegas(tau: float, lam: float | None):
    """tau -> (omega+, omega-) via the TRT magic parameter (Sim.swift:188-193).

    ``lam=None`` gives SRT (omega- == omega+). Computed in double, returned
    rounded to float32 as the reference's ``(Float, Float)``.
    """
    wp = 1.0 / tau
    if lam is None:
        return float(np.float32(wp)), float(np.float32(wp))
    tm = 0.5 + lam / (tau - 0.5)
    return float(np.float32(wp)), float(np.float32(1.0 / tm))


class GPU:
    """One HIP device context with its own stream (Sim.swift:60-78)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().strh_ctx_create(int(device), ctypes.byref(h)))
        self._h = h
        buf = ctypes.create_string_buffer(128)
        check(lib().strh_device_name(h, buf, 128))
        self.name = buf.value.decode()

    def close(self):
        if getattr(self, "_h", None):
            lib().strh_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_gpu = None


def default_gpu() -> GPU:
    global _default_gpu
    if _default_gpu is None:
        _default_gpu = GPU(0)
    return _default_gpu


def flags_array(flags, nx: int, ny: int, nz: int) -> np.ndarray:
    """Raw Cell values, shape (nz, ny, nx), from an array or a vectorised
    closure ``flags(x, y, z)`` evaluated on integer index grids (the Swift
    ``(Int, Int, Int) -> Cell`` closure of Sim.swift:200, applied to all cells
    at once)."""
    if flags is None:
        return np.zeros((nz, ny, nx), dtype=np.uint8)
    if callable(flags):
        z, y, x = np.indices((nz, ny, nx), dtype=np.int64)
        out = np.broadcast_to(np.asarray(flags(x, y, z)), (nz, ny, nx))
        return np.ascontiguousarray(out, dtype=np.uint8)
    a = np.asarray(flags, dtype=np.uint8)
    if a.size != nx * ny * nz:
        raise ValueError("flags must hold nx*ny*nz cells")
    return np.ascontiguousarray(a.reshape(nz, ny, nx))


class Simulation:
    """A D3Q19 lattice on the GPU with the reference's ``Simulation`` surface."""

    def __init__(self, gpu: GPU | None = None, precision="fp32", *, nx: int, ny: int, nz: int,
                 omega: float, omega_minus: float | None = None, lid=(0.0, 0.0, 0.0),
                 uin: float = 0.0, ramp_steps: int = 0, force=(0.0, 0.0, 0.0),
                 c_smago: float = 0.0, wants_forces: bool = False, flags=None):
        self.gpu = gpu if gpu is not None else default_gpu()
        self.precision = Precision(precision)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        fl = flags_array(flags, self.nx, self.ny, self.nz)
        d = LatticeDesc()
        d.precision = self.precision.code
        d.nx, d.ny, d.nz = self.nx, self.ny, self.nz
        d.omega = omega
        d.omega_minus = -1.0 if omega_minus is None else omega_minus
        d.lid[:] = [float(v) for v in lid]
        d.uin = uin
        d.ramp_steps = int(ramp_steps)
        d.force[:] = [float(v) for v in force]
        d.c_smago = c_smago
        d.wants_forces = int(bool(wants_forces))
        h = ctypes.c_void_p()
        check(lib().strh_lattice_create(self.gpu._h, ctypes.byref(d),
                                        fl.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h)))
        self._h = h
        self.flags = fl
        self.wants_forces = bool(wants_forces)
        self._omega = float(np.float32(omega))
        self._omega_minus = float(np.float32(omega if omega_minus is None else omega_minus))
        self._lid = tuple(float(np.float32(v)) for v in lid)
        self._uin = float(np.float32(uin))
        self._ramp = int(ramp_steps)
        self._force = tuple(float(np.float32(v)) for v in force)
        self._c_smago = float(np.float32(c_smago))
        self._inflow_uniform = False
        self._sponge = None

    # ---- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().strh_lattice_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    # ---- mutable public vars (Sim.swift:164-172) --------------------------
    @property
    def omega(self):
        return self._omega

    @omega.setter
    def omega(self, v):
        self._omega = float(np.float32(v))
        check(lib().strh_set_omega(self._h, v, self._omega_minus))

    @property
    def omega_minus(self):
        return self._omega_minus

    @omega_minus.setter
    def omega_minus(self, v):
        self._omega_minus = float(np.float32(v))
        check(lib().strh_set_omega(self._h, self._omega, v))

    @property
    def lid_vel(self):
        return self._lid

    @lid_vel.setter
    def lid_vel(self, v):
        self._lid = tuple(float(np.float32(c)) for c in v)
        check(lib().strh_set_lid_velocity(self._h, *[float(c) for c in v]))

    @property
    def uin_target(self):
        return self._uin

    @uin_target.setter
    def uin_target(self, v):
        self._uin = float(np.float32(v))
        check(lib().strh_set_uin(self._h, v))

    @property
    def ramp_steps(self):
        return self._ramp

    @ramp_steps.setter
    def ramp_steps(self, v):
        self._ramp = int(v)
        check(lib().strh_set_ramp_steps(self._h, int(v)))

    @property
    def force(self):
        return self._force

    @force.setter
    def force(self, v):
        self._force = tuple(float(np.float32(c)) for c in v)
        check(lib().strh_set_force(self._h, *[float(c) for c in v]))

    @property
    def c_smago(self):
        return self._c_smago

    @c_smago.setter
    def c_smago(self, v):
        self._c_smago = float(np.float32(v))
        check(lib().strh_set_c_smago(self._h, v))

    @property
    def inflow_uniform(self):
        return self._inflow_uniform

    @inflow_uniform.setter
    def inflow_uniform(self, v):
        self._inflow_uniform = bool(v)
        check(lib().strh_set_inflow_uniform(self._h, int(bool(v))))

    @property
    def sponge(self):
        return self._sponge

    @sponge.setter
    def sponge(self, v):
        """``(x0, width, tau)`` or ``None`` (Sim.swift:172)."""
        if v is None:
            check(lib().strh_set_sponge(self._h, 0, 0.0, 1.0, 1.0))
        else:
            x0, w, tau = v
            check(lib().strh_set_sponge(self._h, 1, x0, w, tau))
        self._sponge = None if v is None else tuple(float(np.float32(c)) for c in v)

    # ---- setup -------------------------------------------------------------
    def set_solid_fractions(self, eps):
        """Noble-Torczynski solid fractions, one per cell (Sim.swift:413-421)."""
        e = np.ascontiguousarray(eps, dtype=np.float32).reshape(-1)
        if e.size != self.cells:
            raise ValueError("eps must hold one value per cell")
        check(lib().strh_set_solid_fractions(self._h, e.ctypes.data_as(ctypes.c_void_p)))

    def set_porosity(self, alpha):
        """Brinkman porosity per cell (alpha >= 0), or None to remove it.

        New in this implementation; the reference has only solid fractions.
        """
        if alpha is None:
            check(lib().strh_set_porosity(self._h, None))
            return
        a = np.ascontiguousarray(alpha, dtype=np.float32).reshape(-1)
        if a.size != self.cells:
            raise ValueError("alpha must hold one value per cell")
        check(lib().strh_set_porosity(self._h, a.ctypes.data_as(ctypes.c_void_p)))

    # ---- differentiation (no reference counterpart) ----------------------
    @staticmethod
    def _objective(inlet_x, outlet_x, spinup, avg, w_p, w_d, w_f=0.0):
        from ._lib import Objective
        o = Objective()
        o.inlet_x, o.outlet_x = int(inlet_x), int(outlet_x)
        o.w_p, o.w_d, o.w_f = float(w_p), float(w_d), float(w_f)
        o.spinup, o.avg = int(spinup), int(avg)
        return o

    def objective(self, inlet_x: int, outlet_x: int, spinup: int, avg: int,
                  w_p: float = 1.0, w_d: float = 1.0, w_f: float = 0.0) -> float:
        """Advance spinup + avg steps; return the time average over the last avg
        steps of w_p (rho_in - rho_out)/3 + w_d sum alpha J_x + w_f sum ntF_x (FP64)."""
        o = self._objective(inlet_x, outlet_x, spinup, avg, w_p, w_d, w_f)
        J = ctypes.c_double()
        check(lib().strh_objective_eval(self._h, ctypes.byref(o), ctypes.byref(J)))
        return J.value

    def vjp(self, inlet_x: int, outlet_x: int, spinup: int, avg: int, w_p: float = 1.0,
            w_d: float = 1.0, checkpoint_every: int = -1, want_alpha: bool = True,
            w_f: float = 0.0, want_eps: bool = False, inner_every: int | None = None):
        """Discrete adjoint of `objective`: returns (J, dJ/dalpha per cell as
        float64 (nz, ny, nx) or None, dJ/d uin_target, info dict); with
        want_eps, info["grad_eps"] holds dJ/deps (nz, ny, nx) on the cells with
        a solid fraction. Leaves the lattice as `run(spinup + avg)` would.
        checkpoint_every: steps per replayed segment; 0 = ceil(sqrt(T)),
        -1 = as long as memory allows (two levels when one does not fit).
        inner_every: two checkpoint levels (strh_vjp_nested), outer segments of
        checkpoint_every steps replayed once to checkpoint inner segments of
        inner_every steps."""
        from ._lib import VjpInfo
        o = self._objective(inlet_x, outlet_x, spinup, avg, w_p, w_d, w_f)
        J, gu, info = ctypes.c_double(), ctypes.c_double(), VjpInfo()
        ga = np.zeros(self.cells, dtype=np.float64) if want_alpha else None
        ge = np.zeros(self.cells, dtype=np.float64) if want_eps else None
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None
        if inner_every is None:
            check(lib().strh_vjp(self._h, ctypes.byref(o), int(checkpoint_every), ctypes.byref(J),
                                 ptr(ga), ptr(ge), ctypes.byref(gu), ctypes.byref(info)))
        else:
            check(lib().strh_vjp_nested(self._h, ctypes.byref(o), int(checkpoint_every), int(inner_every),
                                        ctypes.byref(J), ptr(ga), ptr(ge), ctypes.byref(gu),
                                        ctypes.byref(info)))
        meta = {k: getattr(info, k) for k, _ in VjpInfo._fields_}
        if want_eps:
            meta["grad_eps"] = ge.reshape(self.nz, self.ny, self.nx)
        return (J.value, ga.reshape(self.nz, self.ny, self.nx) if want_alpha else None,
                gu.value, meta)

    def init_field(self, mode: int, amplitude: float):
        """mode 0 rest, 1 2-D Taylor-Green, 2 3-D TGV (Sim.swift:424-439)."""
        check(lib().strh_init_field(self._h, int(mode), float(amplitude)))

    # ---- stepping ----------------------------------------------------------
    def run(self, steps: int):
        check(lib().strh_run(self._h, int(steps)))

    def run_timed(self, steps: int) -> np.ndarray:
        """run(steps) with a HIP event pair around every step launch: the
        duration of each step in milliseconds (float32, one per step)."""
        ms = np.zeros(int(steps), dtype=np.float32)
        check(lib().strh_run_timed(self._h, int(steps), ms.ctypes.data_as(ctypes.c_void_p)))
        return ms

    @property
    def steps_done(self) -> int:
        v = ctypes.c_int64()
        check(lib().strh_steps_done(self._h, ctypes.byref(v)))
        return v.value

    @property
    def gpu_seconds(self) -> float:
        v = ctypes.c_double()
        check(lib().strh_gpu_seconds(self._h, ctypes.byref(v)))
        return v.value

    # ---- observables -------------------------------------------------------
    def probe_moments(self) -> np.ndarray:
        """(cells, 4) float32 of (ux, uy, uz, rho); even step counts only."""
        out = np.empty((self.cells, 4), dtype=np.float32)
        check(lib().strh_probe_moments(self._h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def probe_force(self, x_range, y_range) -> np.ndarray:
        """Momentum-exchange force over the inclusive box (Sim.swift:359-372)."""
        out = np.zeros(3, dtype=np.float64)
        check(lib().strh_probe_force(self._h, int(x_range[0]), int(x_range[-1]),
                                     int(y_range[0]), int(y_range[-1]),
                                     out.ctypes.data_as(ctypes.c_void_p)))
        return out

    @property
    def state_digest(self) -> str:
        out = (ctypes.c_uint8 * 32)()
        check(lib().strh_state_digest(self._h, out))
        return bytes(out).hex()

    def mass_sum(self) -> float:
        v = ctypes.c_double()
        check(lib().strh_mass_sum(self._h, ctypes.byref(v)))
        return v.value

    def kinetic_energy(self) -> float:
        """FP64 sum over cells of |u|^2 (GPU fixed-order reduction)."""
        v = ctypes.c_double()
        check(lib().strh_kinetic_energy(self._h, ctypes.byref(v)))
        return v.value

    def save_state(self, path: str):
        """Write the DDFs, step count and digest to a state file."""
        check(lib().strh_save_state(self._h, os.fsencode(path)))

    def load_state(self, path: str):
        """Restore a state file written by save_state (same precision and size)."""
        check(lib().strh_load_state(self._h, os.fsencode(path)))

    def dissipation(self) -> float:
        """FP64 sum over fluid cells of 2 nu S:S, S from the non-equilibrium
        momentum flux (GPU fixed-order reduction; even step count)."""
        v = ctypes.c_double()
        check(lib().strh_dissipation(self._h, ctypes.byref(v)))
        return v.value

    # ---- raw state -------------------------------------------------------------
    def download_ddf(self) -> np.ndarray:
        """Canonical DDF image, shape (19, cells), slot-major (i*N + n)."""
        out = np.empty((19, self.cells), dtype=self.precision.dtype)
        check(lib().strh_download_ddf(self._h, out.ctypes.data_as(ctypes.c_void_p), out.nbytes))
        return out

    def upload_ddf(self, a):
        a = np.ascontiguousarray(a).reshape(19, self.cells)
        if a.dtype.itemsize != self.precision.ddf_bytes:
            raise ValueError("DDF image has the wrong element size for this precision")
        check(lib().strh_upload_ddf(self._h, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))

    def download_force(self) -> np.ndarray:
        out = np.empty((self.cells, 4), dtype=np.float32)
        check(lib().strh_download_force(self._h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def snapshot(self):
        check(lib().strh_snapshot(self._h))

    def restore(self):
        check(lib().strh_restore(self._h))


def sha256(data: bytes) -> str:
    """SHA-256 of a host buffer through the library's own implementation."""
    buf = ctypes.create_string_buffer(bytes(data), len(data))
    out = (ctypes.c_uint8 * 32)()
    check(lib().strh_sha256(buf, len(data), out))
    return bytes(out).hex()


def sha256_force_portable(enable: bool) -> bool:
    """Force the portable SHA-256 block function; returns whether SHA-NI is in use."""
    return bool(lib().strh_sha256_force_portable(int(bool(enable))))


__all__ = ["Precision", "Cell", "GPU", "Simulation", "StrouhalError", "trt_omegas",
           "flags_array", "default_gpu", "sha256", "sha256_force_portable"]
