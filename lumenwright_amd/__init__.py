"""lumenwright_amd: Lumenwright's HQ voxel path tracer on AMD Instinct MI355X (gfx950).

A HIP backend behind the reference's Backend contract (crates/lumenwright/src/render/
mod.rs:92-103), reached through the C ABI of include/lw/abi.h. The package holds:

  backend  HipBackend / HipScene: render_hq, the kernel-level trace and render_sample
  frames   host-side frame preparation (sky, camera, Params) and the synthetic frames
  abi      numpy/ctypes mirrors of the scene ABI
  build    the in-tree hipcc build of lumenwright_amd/_lib/liblumenwright_hip.so
"""

from .abi import (EMITTER, HIT, MODEL, PARAMS, QUAD, BoundView, HqFrame, HqScene, VoxelVolume,
                  new_params)
from .backend import BackendError, HipBackend, HipScene, probe

__all__ = ["EMITTER", "HIT", "MODEL", "PARAMS", "QUAD", "BoundView", "HqFrame", "HqScene",
           "VoxelVolume", "new_params", "BackendError", "HipBackend", "HipScene", "probe"]
