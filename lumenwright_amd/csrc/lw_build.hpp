// The device scene builder (lw_build.hip): build_hq's layout (crates/lumenwright/src/scene/
// voxel.rs:305-491, emitter_list :542-607) from an interned voxel volume already in HBM.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "lw/abi.h"

namespace lwb {

// A failure: `invalid` for input the layout cannot hold (ids past the models, a grid or list
// past 2^32 words), otherwise the HIP call that failed and its code.
struct Error {
  bool invalid;
  hipError_t code;
  std::string what;
};

// Device pointers.
struct Input {
  uint32_t dims[3];
  const uint16_t *voxels;        // 4096 model ids per brick
  const uint8_t *brick_occludes;  // per brick, or null: every voxel's model occludes
  const uint8_t *model_flags;     // LW_MODEL_* per model
  const lw_model *models;
  uint32_t models_len;
  const lw_quad *entity_quads;  // n_entity quads in grid space
  uint32_t n_entity;
  uint32_t entity_base;  // the scene quad index of the first entity quad
};

// Buffers hipMalloc'ed by build(); the caller owns them afterwards (hipFree).
struct Output {
  uint32_t *grid = nullptr;
  size_t grid_len = 0;
  lw_emitter *emitters = nullptr;
  size_t emitters_len = 0;
  uint32_t *egrid = nullptr;
  size_t egrid_len = 0;
  uint2 *occupancy = nullptr;  // per stored brick: its 4x4x4 sub-block occupancy
  size_t occupancy_len = 0;
  uint32_t offsets[5] = {0, 0, 0, 0, 0};
  uint64_t bricks = 0, uniform_bricks = 0, buried_bricks = 0, overlay_bricks = 0, emitters_count = 0;
  uint32_t max_cell_list = 0;  // the longest emitter-grid cell list
  double kernel_ms = 0.0;
};

// Runs the layout on `stream` and waits for it. Throws Error; on a throw nothing stays
// allocated.
void build(hipStream_t stream, const Input &in, Output &out);

}  // namespace lwb
