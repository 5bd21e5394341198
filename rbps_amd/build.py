"""Builds the HIP library (librbgpu.so, gfx950) in-tree with hipcc.

Flags:
  -ffp-contract=off                 no FMA contraction: the float expressions are
                                    evaluated as written (reproducible on the CPU)
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division and square root
  -fno-gpu-flush-denormals-to-zero  keep FP32 denormals
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed-FP32 instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librbgpu.so")
SOURCES = [os.path.join(CSRC, "rbgpu_world.hip")]
HEADERS = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))] + [
    os.path.join(ROOT, "include", "rbgpu.h")]

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", LIB + ".tmp", *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building librbgpu.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_oracle(verbose: bool = False) -> None:
    """Builds the CPU checker (oracle/liboracle.so) and, where the reference sources
    are present, the compiled reference (oracle/_ref/libref.so)."""
    odir = os.path.join(ROOT, "oracle")
    targets = ["all"]
    if os.path.isdir("/root/reference/src"):
        targets.append("ref")
    r = subprocess.run(["make", "-C", odir, "-j8", *targets], capture_output=True, text=True)
    if verbose:
        print(r.stdout[-2000:])
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")


if __name__ == "__main__":
    print(build(verbose=True))
    build_oracle(verbose=True)
