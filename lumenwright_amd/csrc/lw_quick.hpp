// Quick mode on the device (lw_quick.hip): the reference's tile-binned software rasterizer
// (crates/lumenwright/src/render/cpu/raster.rs) with its atlas sampling
// (crates/lumenwright/src/render/texture.rs) and the client's sky behind the geometry.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "lw/abi.h"
#include "lw/host.h"
#include "lw_quick_sky.hpp"

namespace lwq {

constexpr uint32_t kMaxMipLevels = 16;

struct Error {
  bool invalid;
  hipError_t code;
  std::string what;
};

// Device pointers, checked by the caller.
struct Input {
  const lw_quick_vertex *verts;
  uint32_t n_verts;
  const uint32_t *indices;  // the opaque, translucent and see-through lists, concatenated
  uint32_t n_opaque, n_translucent, n_see_through;  // triangles per list
  const uint32_t *atlas;    // level 0
  uint32_t atlas_w, atlas_h;
  const uint32_t *mips;     // levels 1..levels
  uint32_t mip_offset[kMaxMipLevels + 1];  // word offset of level l in `mips`
  uint32_t levels, mip_rows;
  const uint32_t *stars;
  uint32_t width, height;
  float view_proj[16];  // column-major
  lw_quick_sky sky;
  lw::DiscSectors sectors;
  const uint2 *log2_fix;  // libm's log2f where it differs from the rounded f64 log2
  uint32_t n_log2_fix;
};

struct Stats {
  uint64_t bin_entries = 0;  // (tile, triangle) pairs
  double kernel_ms = 0.0;    // the device time of the call's launches, inputs resident
};

// Device buffers a backend keeps between calls (grow-only); workspace_free releases them.
struct Workspace;
Workspace *workspace_new();
void workspace_free(Workspace *ws);

// Renders into out (width * height packed RGBA8 words on the device) and waits for it.
void render(hipStream_t stream, const Input &in, Workspace &ws, uint32_t *out, Stats &stats);

}  // namespace lwq
