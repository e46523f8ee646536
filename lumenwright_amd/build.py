"""Builds the gfx950 HIP library in-tree: lumenwright_amd/_lib/liblumenwright_hip.so.

    python -m lumenwright_amd.build

Float semantics are part of the build: no FMA contraction, correctly rounded f32
division and square root, no fast-math, so that each f32 operation rounds as the
reference's CPU build does. Packed-FP32 instruction selection is off.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "_lib" / "liblumenwright_hip.so"
SOURCES = [CSRC / "lw_hip.hip", CSRC / "lw_host.cpp"]
HEADERS = [CSRC / "lw_math.hpp", CSRC / "lw_device.hpp", CSRC / "lw_host.hpp",
           ROOT / "include" / "lw" / "abi.h", ROOT / "include" / "lw" / "host.h"]

ARCH = os.environ.get("LW_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    OUT.parent.mkdir(parents=True, exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), *FLAGS, "-I", str(ROOT / "include"), "-I", str(CSRC), "-x", "hip",
           *[str(s) for s in SOURCES], "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode})")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
