"""strouhal_amd — MI355X-native D3Q19 lattice-Boltzmann solver.

The hot path (the fused AA-pattern collide-and-stream step of the reference's
Sources/StrouhalCore/Kernels.metal:105-355) runs as hand-written HIP kernels
for gfx950 in ``libstrouhal_hip.so``, behind the C ABI of
``include/strouhal_hip.h``. This package is the host-side mirror of the
reference's Swift ``Simulation`` API over that ABI.
"""
from .simulation import (GPU, Cell, Precision, Simulation, StrouhalError, default_gpu,
                         flags_array, sha256, sha256_force_portable, trt_omegas)

__all__ = ["GPU", "Cell", "Precision", "Simulation", "StrouhalError", "default_gpu",
           "flags_array", "sha256", "sha256_force_portable", "trt_omegas"]
