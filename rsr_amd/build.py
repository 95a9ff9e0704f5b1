"""Builds the gfx950 rgl pipeline library rsr_amd/librgl_hip.so with hipcc.

Flags that parity depends on:
  -ffp-contract=off                         no FMA contraction
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE division and sqrt
  -fno-gpu-flush-denormals-to-zero          keep denormals, as SSE does
Packed FP32 instructions are disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "librgl_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RGL_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    src = os.path.join(CSRC, "rgl_hip.hip")
    cmd = [HIPCC, *FLAGS, "-o", LIB + ".tmp", src]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building librgl_hip.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
