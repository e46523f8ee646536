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
SOURCES = [CSRC / "lw_hip.hip", CSRC / "lw_build.hip", CSRC / "lw_quick.hip", CSRC / "lw_host.cpp",
           CSRC / "lw_quick_sky.cpp"]
HEADERS = [CSRC / "lw_math.hpp", CSRC / "lw_device.hpp", CSRC / "lw_host.hpp", CSRC / "lw_build.hpp", CSRC / "lw_sky.hpp", CSRC / "lw_quick_sky.hpp",
           CSRC / "lw_quick.hpp", CSRC / "lw_libm.hpp",
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
    """Compiles each source to an object in parallel (lumenwright_amd/_lib/obj/), then links."""
    OUT.parent.mkdir(parents=True, exist_ok=True)
    objdir = OUT.parent / "obj"
    objdir.mkdir(exist_ok=True)
    inc = ["-I", str(ROOT / "include"), "-I", str(CSRC)]
    compile_flags = [f for f in FLAGS if f != "-shared"]
    jobs = []
    for src in SOURCES:
        obj = objdir / (src.name + ".o")
        cmd = [hipcc(), *compile_flags, *inc, "-x", "hip", "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((src, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                                text=True)))
    failed = []
    for src, obj, proc in jobs:
        out, _ = proc.communicate()
        if proc.returncode != 0:
            sys.stderr.write(out)
            failed.append(src.name)
    if failed:
        raise RuntimeError(f"hipcc failed on {', '.join(failed)}")
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *[str(o) for _, o, _ in jobs],
           "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"link failed ({r.returncode})")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
