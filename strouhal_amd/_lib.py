"""ctypes binding of libstrouhal_hip.so (the C ABI in include/strouhal_hip.h).

The library is the product: there is no CPU fallback. If it is missing this
module raises, naming the build command.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstrouhal_hip.so")

STRH_OK = 0
STATUS_NAMES = {
    0: "STRH_OK", 1: "STRH_ERR_INVALID", 2: "STRH_ERR_NO_DEVICE", 3: "STRH_ERR_DEVICE",
    4: "STRH_ERR_ALLOC", 5: "STRH_ERR_ODD_STEPS", 6: "STRH_ERR_NO_FORCES",
    7: "STRH_ERR_NOT_FINITE", 8: "STRH_ERR_UNSUPPORTED",
}

# Every symbol include/strouhal_hip.h declares.
EXPORTS = [
    "strh_ctx_create", "strh_ctx_destroy", "strh_last_error", "strh_device_name",
    "strh_lattice_create", "strh_lattice_destroy", "strh_lattice_cells", "strh_ddf_bytes",
    "strh_set_omega", "strh_set_lid_velocity", "strh_set_uin", "strh_set_ramp_steps",
    "strh_set_force", "strh_set_c_smago", "strh_set_inflow_uniform", "strh_set_sponge",
    "strh_set_solid_fractions", "strh_set_porosity", "strh_init_field", "strh_run", "strh_run_timed",
    "strh_steps_done",
    "strh_gpu_seconds", "strh_probe_moments", "strh_probe_force", "strh_state_digest",
    "strh_mass_sum", "strh_kinetic_energy", "strh_kinetic_energy_noalloc", "strh_dissipation", "strh_download_ddf", "strh_upload_ddf",
    "strh_download_force", "strh_save_state", "strh_load_state", "strh_snapshot", "strh_restore", "strh_sha256",
    "strh_sha256_force_portable", "strh_batch_create", "strh_batch_destroy",
    "strh_batch_set_probe", "strh_batch_run", "strh_batch_probe_series",
    "strh_batch_set_block_size", "strh_batch_autotune",
    "strh_batch_set_graph_steps", "strh_batch_set_resident", "strh_batch_set_resident_timeout", "strh_batch_gpu_seconds", "strh_objective_eval", "strh_vjp",
    "strh_vjp_nested",
]


class LatticeDesc(ctypes.Structure):
    """strh_lattice_desc: the arguments of Simulation.init (Sim.swift:195-200)."""
    _fields_ = [
        ("precision", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("omega", ctypes.c_float), ("omega_minus", ctypes.c_float),
        ("lid", ctypes.c_float * 3),
        ("uin", ctypes.c_float),
        ("ramp_steps", ctypes.c_int64),
        ("force", ctypes.c_float * 3),
        ("c_smago", ctypes.c_float),
        ("wants_forces", ctypes.c_int32),
    ]


class Objective(ctypes.Structure):
    """strh_objective: time-averaged pressure drop plus Brinkman drag."""
    _fields_ = [
        ("inlet_x", ctypes.c_int32), ("outlet_x", ctypes.c_int32),
        ("w_p", ctypes.c_double), ("w_d", ctypes.c_double),
        ("spinup", ctypes.c_int64), ("avg", ctypes.c_int64),
        ("w_f", ctypes.c_double),
    ]


class VjpInfo(ctypes.Structure):
    """strh_vjp_info: checkpoint plan, timings and workspace of one strh_vjp call."""
    _fields_ = [
        ("checkpoint_every", ctypes.c_int64), ("checkpoints", ctypes.c_int64),
        ("forward_seconds", ctypes.c_double), ("recompute_seconds", ctypes.c_double),
        ("backward_seconds", ctypes.c_double), ("workspace_bytes", ctypes.c_uint64),
        ("outer_every", ctypes.c_int64), ("reverse_pass_seconds", ctypes.c_double),
    ]


class StrouhalError(RuntimeError):
    """Raised for a non-zero status; mirrors Swift's StrouhalError (Sim.swift:127-136)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " or `python strouhal_amd/build.py` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, u32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint32
    f32, sz = ctypes.c_float, ctypes.c_size_t
    sig = {
        "strh_ctx_create": [i32, ctypes.POINTER(vp)],
        "strh_ctx_destroy": [vp],
        "strh_device_name": [vp, ctypes.c_char_p, sz],
        "strh_lattice_create": [vp, ctypes.POINTER(LatticeDesc), vp, ctypes.POINTER(vp)],
        "strh_lattice_destroy": [vp],
        "strh_lattice_cells": [vp, ctypes.POINTER(i64)],
        "strh_ddf_bytes": [vp, ctypes.POINTER(sz)],
        "strh_set_omega": [vp, f32, f32],
        "strh_set_lid_velocity": [vp, f32, f32, f32],
        "strh_set_uin": [vp, f32],
        "strh_set_ramp_steps": [vp, i64],
        "strh_set_force": [vp, f32, f32, f32],
        "strh_set_c_smago": [vp, f32],
        "strh_set_inflow_uniform": [vp, i32],
        "strh_set_sponge": [vp, i32, f32, f32, f32],
        "strh_set_solid_fractions": [vp, vp],
        "strh_set_porosity": [vp, vp],
        "strh_init_field": [vp, u32, f32],
        "strh_run": [vp, i64],
        "strh_run_timed": [vp, i64, vp],
        "strh_steps_done": [vp, ctypes.POINTER(i64)],
        "strh_gpu_seconds": [vp, ctypes.POINTER(ctypes.c_double)],
        "strh_probe_moments": [vp, vp],
        "strh_probe_force": [vp, i32, i32, i32, i32, vp],
        "strh_state_digest": [vp, vp],
        "strh_mass_sum": [vp, ctypes.POINTER(ctypes.c_double)],
        "strh_kinetic_energy": [vp, ctypes.POINTER(ctypes.c_double)],
        "strh_kinetic_energy_noalloc": [vp, ctypes.POINTER(ctypes.c_double)],
        "strh_dissipation": [vp, ctypes.POINTER(ctypes.c_double)],
        "strh_download_ddf": [vp, vp, sz],
        "strh_upload_ddf": [vp, vp, sz],
        "strh_download_force": [vp, vp],
        "strh_save_state": [vp, ctypes.c_char_p],
        "strh_load_state": [vp, ctypes.c_char_p],
        "strh_snapshot": [vp],
        "strh_restore": [vp],
        "strh_sha256": [vp, sz, vp],
        "strh_sha256_force_portable": [i32],
        "strh_batch_create": [vp, ctypes.POINTER(vp), i32, ctypes.POINTER(vp)],
        "strh_batch_destroy": [vp],
        "strh_batch_set_probe": [vp, i32, i64, i64, i64, i32, i32, i32, i32],
        "strh_batch_run": [vp, ctypes.POINTER(i64)],
        "strh_batch_probe_series": [vp, i32, vp, i64, ctypes.POINTER(i64)],
        "strh_batch_set_block_size": [vp, i32],
        "strh_batch_autotune": [vp, ctypes.POINTER(i32)],
        "strh_batch_set_graph_steps": [vp, i32],
        "strh_batch_set_resident": [vp, i32],
        "strh_batch_set_resident_timeout": [vp, ctypes.c_double],
        "strh_batch_gpu_seconds": [vp, ctypes.POINTER(ctypes.c_double)],
        "strh_objective_eval": [vp, ctypes.POINTER(Objective), ctypes.POINTER(ctypes.c_double)],
        "strh_vjp": [vp, ctypes.POINTER(Objective), i64, ctypes.POINTER(ctypes.c_double), vp, vp, vp,
                     ctypes.POINTER(VjpInfo)],
        "strh_vjp_nested": [vp, ctypes.POINTER(Objective), i64, i64, ctypes.POINTER(ctypes.c_double), vp, vp,
                            vp, ctypes.POINTER(VjpInfo)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = i32
    L.strh_last_error.argtypes = []
    L.strh_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def check(rc: int):
    if rc != STRH_OK:
        msg = lib().strh_last_error()
        raise StrouhalError(rc, msg.decode() if msg else "")
