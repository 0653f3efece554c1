"""Builds libsspt.so (HIP for gfx950) in-tree with hipcc.

    python -m seesharp_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsspt.so")
SOURCES = ["api.hip", "kernels.hip", "bvh_build.cpp", "bvh_gpu.hip"]
HEADERS = ["device_math.h", "scene_types.h", "shading.h", "adjoint.h", "kernels.h", "bvh_build.h"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # camera rays, RNG jitter and the watertight triangle test must round like the CPU
    "-ffp-contract=off",
    "-munsafe-fp-atomics",
    # no packed fp32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-Wno-unused-value",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, extra_flags: list[str] | None = None, out: str | None = None) -> str:
    """Compiles the library to `out` (default: the in-tree libsspt.so).
    extra_flags: e.g. ["-DSSPT_TRACE_WAVES=6"] to change the occupancy targets."""
    out = out or OUT
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = out + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, *(extra_flags or []), "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
