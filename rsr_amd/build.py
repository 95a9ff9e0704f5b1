"""Builds the gfx950 rgl pipeline library rsr_amd/librgl_hip.so with hipcc.

Flags that parity depends on:
  -ffp-contract=off                         no FMA contraction
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE division and sqrt
  -fno-gpu-flush-denormals-to-zero          keep denormals, as SSE does
Packed FP32 instructions are disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).

  python -m rsr_amd.build [--programs HEADER] [--out LIB]

--programs adds the programs of HEADER (a header that defines
rgl::UserPrograms = rgl::ProgramList<...>) to the built-in ones.
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


def build(verbose: bool = False, programs: str | None = None, out: str = LIB) -> str:
    """Compiles the library to *out*; *programs* names a header of user
    programs to add to the built-in ones."""
    src = os.path.join(CSRC, "rgl_hip.hip")
    extra = []
    if programs is not None:
        header = os.path.abspath(programs)
        if not os.path.isfile(header):
            raise FileNotFoundError(header)
        extra = [f'-DRGL_USER_PROGRAMS="{header}"']
    cmd = [HIPCC, *FLAGS, *extra, "-o", out + ".tmp", src]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed building {os.path.basename(out)}")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="build librgl_hip.so for gfx950")
    ap.add_argument("--programs", help="header defining rgl::UserPrograms")
    ap.add_argument("--out", default=LIB, help="output library path")
    a = ap.parse_args()
    print(build(verbose=True, programs=a.programs, out=os.path.abspath(a.out)))
