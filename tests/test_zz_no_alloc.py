"""GPU: the allocation-free kinetic energy (strh_kinetic_energy_noalloc)
against strh_kinetic_energy, through the C ABI.

The no-allocation entry point reduces |u|^2 straight from the DDFs in a
static device array; its result must be the same bits as the allocating
entry point's (moments field, then the same fixed-order reduction). Two calls
in a row, the second on a smaller lattice: the static array keeps the first
call's partial sums, and the second must not read any of them.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from strouhal_amd._lib import STRH_OK, StrouhalError, check, lib

pytestmark = pytest.mark.gpu


def _noalloc(sim) -> float:
    out = ctypes.c_double()
    check(lib().strh_kinetic_energy_noalloc(sim._h, ctypes.byref(out)))
    return out.value


def _lattice(gpu, precision, n, walls, force):
    from strouhal_amd import Simulation
    nx, ny, nz = n
    flags = None
    if walls:
        flags = lambda x, y, z: np.where((y == 0) | (y == ny - 1), 1, np.where(x == 0, 2, 0))
    sim = Simulation(gpu, precision, nx=nx, ny=ny, nz=nz, omega=1.7, omega_minus=1.4, force=force,
                     lid=(0.0, 0.03, 0.0), flags=flags)
    sim.init_field(2, 0.05)
    sim.run(6)
    return sim


@pytest.mark.parametrize("precision", ["fp32", "fp16s", "bf16"])
def test_noalloc_kinetic_energy_equals_the_allocating_entry_point(gpu, precision):
    cases = [((200, 160, 130), True, (1e-5, -2e-5, 0.0)),    # 4.2M cells: 1,017 partial sums
             ((37, 21, 5), False, (0.0, 0.0, 0.0))]          # then a small lattice: 1 partial
    for n, walls, force in cases:
        sim = _lattice(gpu, precision, n, walls, force)
        b = _noalloc(sim)
        a = sim.kinetic_energy()
        assert np.isfinite(a) and a > 0.0
        assert np.float64(b).view(np.uint64) == np.float64(a).view(np.uint64), (n, a, b)
        assert _noalloc(sim) == b                            # repeatable
        sim.close()


def test_noalloc_kinetic_energy_refuses_odd_step_counts(gpu):
    sim = _lattice(gpu, "fp32", (16, 8, 4), False, (0.0, 0.0, 0.0))
    sim.run(1)
    out = ctypes.c_double()
    rc = lib().strh_kinetic_energy_noalloc(sim._h, ctypes.byref(out))
    assert rc != STRH_OK
    with pytest.raises(StrouhalError):
        check(rc)
    sim.close()
