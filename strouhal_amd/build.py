"""Build the HIP solver library in-tree: strouhal_amd/libstrouhal_hip.so.

Device code (the csrc/*.hip translation units) is compiled by hipcc for
gfx950, in parallel; the host runtime and SHA-256 are plain C++ (g++) linked
against the HIP runtime. All code is
built with -ffp-contract=off (results round after every operation except the
explicit fused multiply-adds) and packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libstrouhal_hip.so")
# The streaming-ceiling measurement kernel (profile_workload.py --stream-ceiling)
# lives in a library of its own, outside the solver's C ABI.
STREAM_LIB = os.path.join(HERE, "libstrouhal_stream.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fvisibility=hidden", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# Device translation units (the storage formats of the step kernels each have one).
DEVICE_UNITS = ["step_fp32", "step_fp16", "step_bf16", "adjoint", "batch", "resident", "observe"]
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


def build(verbose: bool = False, jobs: int | None = None) -> str:
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    os.makedirs(BUILD, exist_ok=True)
    inc = f"-I{os.path.join(ROOT, 'include')}"
    cmds, objs = [], []
    for u in DEVICE_UNITS:
        o = os.path.join(BUILD, u + ".o")
        cmds.append([hipcc, *DEVICE_FLAGS, inc, "-c", os.path.join(CSRC, u + ".hip"), "-o", o])
        objs.append(o)
    stream_obj = os.path.join(BUILD, "stream_ceiling.o")
    cmds.append([hipcc, *DEVICE_FLAGS, "-c", os.path.join(CSRC, "stream_ceiling.hip"), "-o", stream_obj])
    for u in ("runtime", "sha256"):
        o = os.path.join(BUILD, u + ".o")
        cmds.append(["g++", *HOST_FLAGS, inc, "-c", os.path.join(CSRC, u + ".cpp"), "-o", o])
        objs.append(o)
    if jobs is None:
        jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp, "-lpthread"], verbose)
    os.replace(tmp, LIB)
    tmp = STREAM_LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", stream_obj, "-o", tmp], verbose)
    os.replace(tmp, STREAM_LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
