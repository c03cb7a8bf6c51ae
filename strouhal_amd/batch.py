"""Multi-lattice batches over the HIP C ABI (strh_batch_*).

A batch advances several lattices together, each with its own step count,
ramp and momentum-exchange probe schedule held on the device: lattices one
cell deep are held in LDS for up to 256 steps per launch (resident launches,
where they measure faster), others stream through memory with one launch per
step (replayed as HIP graphs). It replaces the
sequential rung loop of the reference's Ladder.credibility / Ladder.measure
(Sources/StrouhalCore/Ladder.swift:54-121); every batched lattice evolves bit
for bit as it does when run alone.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib
from .simulation import Simulation


class Batch:
    def __init__(self, sims: list[Simulation]):
        if not sims:
            raise ValueError("a batch needs at least one lattice")
        self.sims = list(sims)
        arr = (ctypes.c_void_p * len(sims))(*[s._h for s in sims])
        h = ctypes.c_void_p()
        check(lib().strh_batch_create(sims[0].gpu._h, arr, len(sims), ctypes.byref(h)))
        self._h = h
        self._counts = [0] * len(sims)

    def close(self):
        if getattr(self, "_h", None):
            lib().strh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_probe(self, index: int, first_step: int, period: int, count: int, x_range, y_range):
        """Force samples on odd steps first_step + j*period, j < count (box inclusive)."""
        check(lib().strh_batch_set_probe(self._h, int(index), int(first_step), int(period),
                                         int(count), int(x_range[0]), int(x_range[-1]),
                                         int(y_range[0]), int(y_range[-1])))
        self._counts[index] = int(count)

    def run(self, steps):
        """Advance lattice i by steps[i] (an int applies to all)."""
        if np.isscalar(steps):
            steps = [int(steps)] * len(self.sims)
        arr = (ctypes.c_int64 * len(self.sims))(*[int(s) for s in steps])
        check(lib().strh_batch_run(self._h, arr))

    def probe_series(self, index: int) -> np.ndarray:
        """(samples, 3) float64: the FP64 box sums (Fx, Fy, Fz) taken so far."""
        cap = max(self._counts[index], 0)
        out = np.zeros((max(cap, 1), 3), dtype=np.float64)
        taken = ctypes.c_int64()
        check(lib().strh_batch_probe_series(self._h, int(index), out.ctypes.data_as(ctypes.c_void_p),
                                            cap, ctypes.byref(taken)))
        return out[:taken.value].copy()

    def set_block_size(self, bs: int):
        check(lib().strh_batch_set_block_size(self._h, int(bs)))

    def autotune(self) -> int:
        v = ctypes.c_int()
        check(lib().strh_batch_autotune(self._h, ctypes.byref(v)))
        return v.value

    def set_graph_steps(self, steps: int):
        check(lib().strh_batch_set_graph_steps(self._h, int(steps)))

    def set_resident(self, mode: int):
        """0: stream every DDF through memory each step; 1: hold lattices one
        cell deep in LDS across steps (resident launches) where the active
        lattices fit on chip in one launch; 2: resident launches everywhere,
        time-sliced groups where needed (an error for lattices that do not
        qualify); 3 (default): as 1, and where the active lattices would need
        time-sliced groups, time both kinds of launch over the segment's first
        steps and keep the faster."""
        check(lib().strh_batch_set_resident(self._h, int(mode)))

    def set_resident_timeout(self, seconds: float):
        """How long a resident workgroup waits for a neighbour's exchange word
        before the run fails (default 1 s)."""
        check(lib().strh_batch_set_resident_timeout(self._h, float(seconds)))

    @property
    def gpu_seconds(self) -> float:
        v = ctypes.c_double()
        check(lib().strh_batch_gpu_seconds(self._h, ctypes.byref(v)))
        return v.value
