"""Builds the HIP extension libconvector_amd.so for gfx950 (in-tree).

Every kernel is compiled with -ffp-contract=off (products are fused only
where the source writes fmaf) and correctly rounded f32 division/sqrt, so
the device arithmetic matches the reference's AVX arithmetic bit for bit.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libconvector_amd.so")
SOURCES = ["api.hip", "render.hip", "kernels.hip", "build.hip", "denoise.hip", "wide.hip"]
HEADERS = ["device_math.hpp", "internal.hpp", "host.hpp"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags(arch: str = "gfx950"):
    return [
        "-O3", "-std=c++17", "-fPIC", "-shared",
        f"--offload-arch={arch}",
        "-ffp-contract=off", "-fno-fast-math",
        "-fhip-fp32-correctly-rounded-divide-sqrt",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-result",
    ]


def build(verbose: bool = False, extra=None) -> str:
    """Compiles every HIP source into libconvector_amd.so (always rebuilds)."""
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + flags() + (extra or []) + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
