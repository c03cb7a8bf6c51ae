"""ctypes wrapper of the CPU oracle — TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this module, and only as the checker. The product (``strouhal_amd``)
never imports it.

``OracleSim`` mirrors the surface of the reference's ``Simulation``
(Sources/StrouhalCore/Sim.swift:153-459) so parity tests can drive the HIP
library and the oracle with the same calls:

* ``run(steps)``               Sim.swift:288-331
* ``probe_moments()``          Sim.swift:334-353  (Kernels.metal:359-380)
* ``probe_force(xr, yr)``      Sim.swift:359-372
* ``init_field(mode, amp)``    Sim.swift:424-439  (Kernels.metal:393-453)
* ``set_solid_fractions(eps)`` Sim.swift:413-421
* ``mass_sum()``               Sim.swift:447-458
* ``state_digest``             Sim.swift:441-444  (hashlib SHA-256 of the raw DDF bytes)
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libstrouhal_oracle.so")
_lib = None

ARITH = {"f64": 0, "f32": 1}
STORAGE = {"f64": 0, "f32": 1, "f16": 2, "bf16": 3}
STORAGE_DTYPE = {"f64": np.float64, "f32": np.float32, "f16": np.uint16, "bf16": np.uint16}


class HostParams(ctypes.Structure):
    """Host-side solver state, Sim.swift:164-172."""
    _fields_ = [
        ("omega", ctypes.c_double), ("omega_minus", ctypes.c_double),
        ("lid", ctypes.c_double * 3),
        ("uin_target", ctypes.c_double),
        ("ramp_steps", ctypes.c_int64),
        ("force", ctypes.c_double * 3),
        ("c_smago", ctypes.c_double),
        ("inflow_uniform", ctypes.c_int32),
        ("sponge_on", ctypes.c_int32),
        ("sponge_x0", ctypes.c_double), ("sponge_width", ctypes.c_double),
        ("sponge_tau", ctypes.c_double),
    ]


class Objective(ctypes.Structure):
    """The differentiable channel objective (oracle/strouhal_oracle.cpp `Objective`)."""
    _fields_ = [
        ("inlet_x", ctypes.c_int32), ("outlet_x", ctypes.c_int32),
        ("w_p", ctypes.c_double), ("w_d", ctypes.c_double),
        ("spinup", ctypes.c_int64), ("avg", ctypes.c_int64),
        ("w_f", ctypes.c_double),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i32, i64, dp, fp = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float))
        L.orc_create.restype = vp
        L.orc_create.argtypes = [i32, i32, i32, i32, i32, ctypes.c_void_p, i32]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_host_params.argtypes = [vp, ctypes.POINTER(HostParams)]
        L.orc_set_threads.argtypes = [vp, i32]
        L.orc_set_solid_fractions.argtypes = [vp, ctypes.c_void_p]
        L.orc_set_porosity.argtypes = [vp, ctypes.c_void_p]
        L.orc_objective.argtypes = [vp, ctypes.POINTER(Objective)]
        L.orc_objective.restype = ctypes.c_double
        L.orc_vjp.argtypes = [vp, ctypes.POINTER(Objective), ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_void_p, i64]
        L.orc_vjp.restype = ctypes.c_double
        L.orc_init_field.argtypes = [vp, i32, ctypes.c_double]
        L.orc_run.argtypes = [vp, i64]
        L.orc_steps_done.argtypes = [vp]
        L.orc_steps_done.restype = i64
        L.orc_probe_moments.argtypes = [vp, ctypes.c_void_p]
        L.orc_probe_moments.restype = i32
        L.orc_probe_force.argtypes = [vp, i32, i32, i32, i32, ctypes.c_void_p]
        L.orc_probe_force.restype = i32
        L.orc_force_field.argtypes = [vp, ctypes.c_void_p]
        L.orc_mass_sum.argtypes = [vp]
        L.orc_mass_sum.restype = ctypes.c_double
        L.orc_ddf_bytes.argtypes = [vp]
        L.orc_ddf_bytes.restype = ctypes.c_size_t
        L.orc_download_ddf.argtypes = [vp, ctypes.c_void_p]
        L.orc_upload_ddf.argtypes = [vp, ctypes.c_void_p]
        L.orc_masks.argtypes = [vp, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_collide_cell.argtypes = [vp, ctypes.c_void_p, ctypes.c_double, i32,
                                       ctypes.c_void_p, dp, ctypes.c_void_p]
        L.orc_f32_to_f16_array.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.orc_f32_to_bf16_array.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.orc_f16_to_f32_array.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def f32_to_f16(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    out = np.empty(a.shape, dtype=np.uint16)
    lib().orc_f32_to_f16_array(_ptr(a), _ptr(out), a.size)
    return out


def f32_to_bf16(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    out = np.empty(a.shape, dtype=np.uint16)
    lib().orc_f32_to_bf16_array(_ptr(a), _ptr(out), a.size)
    return out


def f16_to_f32(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint16)
    out = np.empty(a.shape, dtype=np.float32)
    lib().orc_f16_to_f32_array(_ptr(a), _ptr(out), a.size)
    return out


class OracleSim:
    """CPU oracle lattice with the reference ``Simulation`` surface.

    ``flags`` is a uint8 array of shape (nz, ny, nx) holding raw ``Cell``
    values (Sim.swift:46-52). ``arith`` is "f64" or "f32"; ``storage`` is
    "f64" (with f64 arithmetic) or "f32" / "f16" / "bf16".
    """

    def __init__(self, nx, ny, nz, omega, omega_minus=None, *, arith="f32",
                 storage="f32", lid=(0.0, 0.0, 0.0), uin=0.0, ramp_steps=0,
                 force=(0.0, 0.0, 0.0), c_smago=0.0, wants_forces=False,
                 flags=None, threads=1):
        if arith == "f64" and storage != "f64":
            raise ValueError("the FP64 oracle stores f64")
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.cells = self.nx * self.ny * self.nz
        self.arith, self.storage = arith, storage
        if flags is None:
            flags = np.zeros((self.nz, self.ny, self.nx), dtype=np.uint8)
        flags = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        if flags.size != self.cells:
            raise ValueError("flags must hold nx*ny*nz cells")
        self.flags = flags.reshape(self.nz, self.ny, self.nx)
        self._h = lib().orc_create(ARITH[arith], STORAGE[storage], self.nx, self.ny,
                                   self.nz, _ptr(flags), int(bool(wants_forces)))
        if not self._h:
            raise ValueError("oracle lattice creation failed")
        self.wants_forces = bool(wants_forces)
        self.omega = float(omega)
        self.omega_minus = float(omega if omega_minus is None else omega_minus)
        self.lid_vel = tuple(float(v) for v in lid)
        self.uin_target = float(uin)
        self.ramp_steps = int(ramp_steps)
        self.force = tuple(float(v) for v in force)
        self.c_smago = float(c_smago)
        self.inflow_uniform = False
        self.sponge = None           # (x0, width, tau) or None
        lib().orc_set_threads(self._h, int(threads))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    def set_threads(self, t: int):
        lib().orc_set_threads(self._h, int(t))

    def _push(self):
        hp = HostParams()
        hp.omega, hp.omega_minus = self.omega, self.omega_minus
        hp.lid[:] = list(self.lid_vel)
        hp.uin_target = self.uin_target
        hp.ramp_steps = self.ramp_steps
        hp.force[:] = list(self.force)
        hp.c_smago = self.c_smago
        hp.inflow_uniform = int(bool(self.inflow_uniform))
        if self.sponge is not None:
            hp.sponge_on = 1
            hp.sponge_x0, hp.sponge_width, hp.sponge_tau = (float(v) for v in self.sponge)
        lib().orc_set_host_params(self._h, ctypes.byref(hp))

    @property
    def steps_done(self) -> int:
        return int(lib().orc_steps_done(self._h))

    def set_solid_fractions(self, eps):
        eps = np.ascontiguousarray(eps, dtype=np.float32).reshape(-1)
        if eps.size != self.cells:
            raise ValueError("eps must hold nx*ny*nz values")
        lib().orc_set_solid_fractions(self._h, _ptr(eps))

    def set_porosity(self, alpha):
        """Brinkman porosity alpha per cell (drag force -alpha rho u)."""
        a = np.ascontiguousarray(alpha, dtype=np.float32).reshape(-1)
        if a.size != self.cells:
            raise ValueError("alpha must hold nx*ny*nz values")
        lib().orc_set_porosity(self._h, _ptr(a))

    @staticmethod
    def _objective(inlet_x, outlet_x, w_p, w_d, spinup, avg, w_f=0.0):
        o = Objective()
        o.inlet_x, o.outlet_x, o.w_p, o.w_d = int(inlet_x), int(outlet_x), float(w_p), float(w_d)
        o.spinup, o.avg, o.w_f = int(spinup), int(avg), float(w_f)
        return o

    def objective(self, inlet_x, outlet_x, spinup, avg, w_p=1.0, w_d=1.0, w_f=0.0) -> float:
        """Run spinup + avg steps (FP64 only) and return the time-averaged objective."""
        self._push()
        o = self._objective(inlet_x, outlet_x, w_p, w_d, spinup, avg, w_f)
        return float(lib().orc_objective(self._h, ctypes.byref(o)))

    def vjp(self, inlet_x, outlet_x, spinup, avg, w_p=1.0, w_d=1.0, checkpoint_every=0, w_f=0.0,
            with_eps=False):
        """FP64 discrete adjoint: (J, dJ/dalpha per cell, dJ/d uin_target), or
        (J, dJ/dalpha, dJ/deps, dJ/d uin_target) with with_eps (dJ/deps on the
        cells whose solid fraction is non-zero, 0 elsewhere).

        Replays segments of checkpoint_every steps (0: ceil(sqrt(T))) from DDF
        checkpoints; the result does not depend on the interval."""
        self._push()
        o = self._objective(inlet_x, outlet_x, w_p, w_d, spinup, avg, w_f)
        ga = np.zeros(self.cells, dtype=np.float64)
        ge = np.zeros(self.cells, dtype=np.float64)
        gu = np.zeros(1, dtype=np.float64)
        J = float(lib().orc_vjp(self._h, ctypes.byref(o), _ptr(ga), _ptr(ge), _ptr(gu),
                                int(checkpoint_every)))
        if with_eps:
            return J, ga, ge, float(gu[0])
        return J, ga, float(gu[0])

    def init_field(self, mode: int, amplitude: float):
        self._push()
        lib().orc_init_field(self._h, int(mode), float(amplitude))

    def run(self, steps: int):
        self._push()
        lib().orc_run(self._h, int(steps))

    def probe_moments(self) -> np.ndarray:
        """(cells, 4) float64 array of (ux, uy, uz, rho) in the oracle's arithmetic."""
        self._push()
        out = np.empty((self.cells, 4), dtype=np.float64)
        if lib().orc_probe_moments(self._h, _ptr(out)) != 0:
            raise RuntimeError("moments probe requires an even step count")
        return out

    def probe_force(self, x_range, y_range) -> np.ndarray:
        self._push()
        out = np.zeros(3, dtype=np.float64)
        rc = lib().orc_probe_force(self._h, int(x_range[0]), int(x_range[1]),
                                   int(y_range[0]), int(y_range[1]), _ptr(out))
        if rc != 0:
            raise RuntimeError("force probe requires an even step count and wants_forces")
        return out

    def force_field(self) -> np.ndarray:
        out = np.zeros((self.cells, 4), dtype=np.float64)
        lib().orc_force_field(self._h, _ptr(out))
        return out

    def mass_sum(self) -> float:
        return float(lib().orc_mass_sum(self._h))

    def download_ddf(self) -> np.ndarray:
        """Raw DDF storage, shape (19, cells), in slot-major order i*N + n."""
        out = np.empty((19, self.cells), dtype=STORAGE_DTYPE[self.storage])
        lib().orc_download_ddf(self._h, _ptr(out))
        return out

    def upload_ddf(self, a):
        a = np.ascontiguousarray(a, dtype=STORAGE_DTYPE[self.storage]).reshape(19, self.cells)
        lib().orc_upload_ddf(self._h, _ptr(a))

    def ddf_float(self) -> np.ndarray:
        """DDFs widened to float64, shape (19, cells)."""
        raw = self.download_ddf()
        if self.storage == "f16":
            return f16_to_f32(raw).astype(np.float64)
        if self.storage == "bf16":
            return (raw.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        return raw.astype(np.float64)

    @property
    def state_digest(self) -> str:
        return hashlib.sha256(self.download_ddf().tobytes()).hexdigest()

    def masks(self):
        s = np.empty(self.cells, dtype=np.uint32)
        l = np.empty(self.cells, dtype=np.uint32)
        lib().orc_masks(self._h, _ptr(s), _ptr(l))
        return s, l

    def collide_cell(self, fin, eps=0.0, x=0):
        """Apply collide (Kernels.metal:105-228) to one cell's 19 DDFs."""
        self._push()
        fin = np.ascontiguousarray(fin, dtype=np.float64)
        fout = np.empty(19, dtype=np.float64)
        rho = ctypes.c_double()
        ntf = np.zeros(3, dtype=np.float64)
        lib().orc_collide_cell(self._h, _ptr(fin), float(eps), int(x), _ptr(fout),
                               ctypes.byref(rho), _ptr(ntf))
        return fout, rho.value, ntf
