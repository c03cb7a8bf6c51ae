"""Build the HIP solver library in-tree: strouhal_amd/libstrouhal_hip.so.

Device code (kernels.hip) is compiled by hipcc for gfx950; the host runtime
and SHA-256 are plain C++ (g++) linked against the HIP runtime. All code is
built with -ffp-contract=off (results round after every operation except the
explicit fused multiply-adds) and packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libstrouhal_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fvisibility=hidden", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HOST_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fvisibility=hidden",
    "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", "-Wall", "-Wextra", "-Wno-unused-result",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:2])} ... ({r.returncode})")
    return r


def build(verbose: bool = False) -> str:
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    os.makedirs(BUILD, exist_ok=True)
    inc = f"-I{os.path.join(ROOT, 'include')}"
    kobj = os.path.join(BUILD, "kernels.o")
    robj = os.path.join(BUILD, "runtime.o")
    sobj = os.path.join(BUILD, "sha256.o")
    _run([hipcc, *DEVICE_FLAGS, inc, "-c", os.path.join(CSRC, "kernels.hip"), "-o", kobj], verbose)
    _run(["g++", *HOST_FLAGS, inc, "-c", os.path.join(CSRC, "runtime.cpp"), "-o", robj], verbose)
    _run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, "sha256.cpp"), "-o", sobj], verbose)
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", kobj, robj, sobj, "-o", tmp,
          "-lpthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
