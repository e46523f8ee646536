// ORACLE -- test infrastructure only. A restatement of build_hq's layout,
// crates/lumenwright/src/scene/voxel.rs:305-491, from the interned voxel volume: the top-level
// entries (EMPTY, buried and UNIFORM bricks collapsed, :305-342), the entity overlay
// (:344-425), the grid buffer (:427-446) and emitter_list (:542-607).
//
// One order is this file's own: the reference appends an overlay brick's cell lists in the
// iteration order of a HashMap (voxel.rs:418-424), which is unspecified; here they follow in
// cell order. Each cell's (start, count) still names the same list of quads, in the same order.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <vector>

namespace lwo_build {

constexpr uint32_t EMPTY = 0xFFFFFFFFu, UNIFORM = 0x80000000u;
constexpr uint32_t BRICK_WORDS = 2048, OV_CELL_WORDS = 8192, GRID_HEADER = 8;
constexpr uint8_t OCCLUDES = 1, EMITS = 2;
constexpr int32_t EMITTER_CELL = 10;  // voxel.rs:532

struct Model { uint32_t quad_offset, quad_count, flags, pad; };
struct Quad { float p0[4], eu[4], ev[4], uv[4], uvt[4]; };
struct Emitter { float pos[4]; uint32_t faces[4]; };

struct Volume {
  uint32_t dims[3];
  const uint16_t *voxels;          // 4096 per brick, brick ((by*dz + bz)*dx + bx)
  const uint8_t *brick_occludes;   // per brick, or null: from the model flags
  const Model *models;
  size_t models_len;
  const uint8_t *model_flags;
  size_t quads_len;                // the models' quads (entity quads follow them)
  const Quad *entity_quads;
  size_t entity_quads_len;
  uint32_t see_through_count;
};

struct Built {
  std::vector<uint32_t> grid;
  std::vector<Emitter> emitters;
  std::vector<uint32_t> emitter_grid;
  uint32_t offsets[5];
  uint32_t see_through[2];
  uint32_t emitter_count;
  uint64_t bricks = 0, uniform_bricks = 0, buried_bricks = 0, overlay_bricks = 0;
};

static inline int32_t as_i32(float x) {  // Rust `as i32`: saturating, NaN -> 0
  if (std::isnan(x)) return 0;
  if (x >= 2147483648.0f) return INT32_MAX;
  if (x <= -2147483648.0f) return INT32_MIN;
  return (int32_t)x;
}
static inline float fmin_r(float a, float b) {  // f32::min: the non-NaN operand
  if (std::isnan(a)) return b;
  if (std::isnan(b)) return a;
  return b < a ? b : a;
}
static inline float fmax_r(float a, float b) {
  if (std::isnan(a)) return b;
  if (std::isnan(b)) return a;
  return b > a ? b : a;
}

// emitter_list (voxel.rs:542-607).
static void emitter_list(const Model *models, const std::vector<std::pair<std::array<int32_t, 3>, uint32_t>> &sites,
                         std::vector<Emitter> &emitters, std::vector<uint32_t> &grid) {
  emitters.clear();
  grid.clear();
  if (sites.empty()) {
    grid.assign(GRID_HEADER, 0u);
    grid[0] = (uint32_t)EMITTER_CELL;
    emitters.push_back(Emitter{});
    return;
  }
  for (auto &s : sites) {
    const Model &m = models[s.second];
    Emitter e{};
    for (int a = 0; a < 3; a++) e.pos[a] = (float)s.first[a];
    std::memcpy(&e.pos[3], &m.quad_offset, 4);
    e.faces[0] = m.quad_count;
    emitters.push_back(e);
  }
  auto cell_of = [](int32_t v) {  // i32::div_euclid
    int32_t q = v / EMITTER_CELL;
    if (v % EMITTER_CELL < 0) q -= 1;
    return q;
  };
  int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  for (auto &s : sites)
    for (int a = 0; a < 3; a++) {
      lo[a] = std::min(lo[a], cell_of(s.first[a]));
      hi[a] = std::max(hi[a], cell_of(s.first[a]));
    }
  int32_t offset[3], size[3];
  for (int a = 0; a < 3; a++) {
    offset[a] = lo[a] - 1;
    size[a] = hi[a] - lo[a] + 3;
  }
  std::vector<std::vector<uint32_t>> cells((size_t)size[0] * size[1] * size[2]);
  for (size_t i = 0; i < sites.size(); i++) {
    int32_t g[3];
    for (int a = 0; a < 3; a++) g[a] = cell_of(sites[i].first[a]);
    for (int dy = -1; dy <= 1; dy++)
      for (int dz = -1; dz <= 1; dz++)
        for (int dx = -1; dx <= 1; dx++) {
          const int64_t l0 = g[0] + dx - offset[0], l1 = g[1] + dy - offset[1], l2 = g[2] + dz - offset[2];
          cells[(size_t)((l1 * size[2] + l2) * size[0] + l0)].push_back((uint32_t)i);
        }
  }
  grid = {(uint32_t)EMITTER_CELL, (uint32_t)size[0], (uint32_t)size[1], (uint32_t)size[2],
          (uint32_t)offset[0], (uint32_t)offset[1], (uint32_t)offset[2], 0u};
  std::vector<uint32_t> list;
  for (auto &c : cells) {
    grid.push_back((uint32_t)list.size());
    grid.push_back((uint32_t)c.size());
    list.insert(list.end(), c.begin(), c.end());
  }
  grid.insert(grid.end(), list.begin(), list.end());
}

static Built build(const Volume &v) {
  const int64_t dx = v.dims[0], dy = v.dims[1], dz = v.dims[2];
  const size_t nb = (size_t)(dx * dy * dz);
  for (size_t i = 0; i < nb * 4096; i++)
    if (v.voxels[i] >= v.models_len) throw std::invalid_argument("voxel with an unknown model");
  auto brick_xyz = [&](size_t i, int32_t &bx, int32_t &by, int32_t &bz) {
    bx = (int32_t)(i % dx);
    by = (int32_t)(i / (dx * dz));
    bz = (int32_t)((i / dx) % dz);
  };
  auto brick_index = [&](int32_t bx, int32_t by, int32_t bz) -> int64_t {
    if (bx < 0 || by < 0 || bz < 0 || bx >= dx || by >= dy || bz >= dz) return -1;
    return ((int64_t)by * dz + bz) * dx + bx;
  };
  // voxel.rs:119-130: per section, every block occludes
  std::vector<uint8_t> all_occluding(nb, 0);
  for (size_t b = 0; b < nb; b++) {
    if (v.brick_occludes) {
      all_occluding[b] = v.brick_occludes[b] ? 1 : 0;
      continue;
    }
    bool all = true;
    for (uint32_t k = 0; k < 4096 && all; k++) all = (v.model_flags[v.voxels[b * 4096 + k]] & OCCLUDES) != 0;
    all_occluding[b] = all ? 1 : 0;
  }
  Built out;
  std::vector<uint32_t> top(nb, EMPTY), bricks;
  std::vector<std::pair<std::array<int32_t, 3>, uint32_t>> sites;
  for (size_t b = 0; b < nb; b++) {
    const uint16_t *ids = v.voxels + b * 4096;
    bool any = false;
    for (uint32_t k = 0; k < 4096; k++) any = any || ids[k] != 0;
    if (!any) continue;  // voxel.rs:290-292: no instance keys, the entry stays EMPTY
    int32_t bx, by, bz;
    brick_xyz(b, bx, by, bz);
    // :307-312
    bool buried = all_occluding[b] != 0;
    static const int32_t N[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (auto &n : N) {
      const int64_t j = brick_index(bx + n[0], by + n[1], bz + n[2]);
      if (j < 0 || !all_occluding[j]) buried = false;
    }
    // :315-324: every emitting block is an emitter, buried or not
    for (uint32_t local = 0; local < 4096; local++)
      if (v.model_flags[ids[local]] & EMITS)
        sites.push_back({{bx * 16 + (int32_t)(local & 15), by * 16 + (int32_t)(local >> 8),
                          bz * 16 + (int32_t)((local >> 4) & 15)},
                         ids[local]});
    if (buried) {  // :325-329
      out.buried_bricks++;
      uint32_t first = 1;
      for (uint32_t k = 0; k < 4096; k++)
        if (ids[k] != 0) {
          first = ids[k];
          break;
        }
      top[b] = UNIFORM | first;
      continue;
    }
    bool uniform = true;  // :330-334
    for (uint32_t k = 1; k < 4096 && uniform; k++) uniform = ids[k] == ids[0];
    if (uniform) {
      out.uniform_bricks++;
      top[b] = UNIFORM | ids[0];
      continue;
    }
    top[b] = (uint32_t)(bricks.size() / BRICK_WORDS);  // :335-341
    for (uint32_t k = 0; k < 4096; k += 2) bricks.push_back((uint32_t)ids[k] | ((uint32_t)ids[k + 1] << 16));
    out.bricks++;
  }

  // The entity overlay (:344-425): every voxel an overlay quad's bounds (grown by 1e-3)
  // touch lists the quad.
  const uint32_t entity_base = (uint32_t)v.quads_len;
  const int32_t size[3] = {(int32_t)dx * 16, (int32_t)dy * 16, (int32_t)dz * 16};
  std::map<size_t, std::map<uint32_t, std::vector<uint32_t>>> cells;  // brick -> cell -> quads
  for (size_t i = 0; i < v.entity_quads_len; i++) {
    const Quad &q = v.entity_quads[i];
    float c[4][3];
    for (int a = 0; a < 3; a++) {
      c[0][a] = q.p0[a];
      c[1][a] = q.p0[a] + q.ev[a];
      c[2][a] = (q.p0[a] + q.eu[a]) + q.ev[a];
      c[3][a] = q.p0[a] + q.eu[a];
    }
    int32_t l[3], h[3];
    for (int a = 0; a < 3; a++) {
      float lo = 3.40282347e38f, hi = -3.40282347e38f;
      for (int k = 0; k < 4; k++) {
        lo = fmin_r(lo, c[k][a]);
        hi = fmax_r(hi, c[k][a]);
      }
      lo = lo - 1e-3f;
      hi = hi + 1e-3f;
      l[a] = std::max(as_i32(std::floor(lo)), 0);
      h[a] = std::min(as_i32(std::floor(hi)), size[a] - 1);
    }
    for (int32_t y = l[1]; y <= h[1]; y++)
      for (int32_t z = l[2]; z <= h[2]; z++)
        for (int32_t x = l[0]; x <= h[0]; x++) {
          const size_t b = (size_t)brick_index(x >> 4, y >> 4, z >> 4);
          const uint32_t local = (uint32_t)(((y & 15) * 16 + (z & 15)) * 16 + (x & 15));
          cells[b][local].push_back(entity_base + (uint32_t)i);
        }
  }
  std::vector<uint32_t> ov_top(nb, EMPTY), ov_cells, ov_quads;
  for (auto &bc : cells) {  // bricks ascending (:413-416)
    ov_top[bc.first] = (uint32_t)(ov_cells.size() / OV_CELL_WORDS);
    const size_t start = ov_cells.size();
    ov_cells.resize(start + OV_CELL_WORDS, 0u);
    for (auto &cl : bc.second) {
      ov_cells[start + cl.first * 2] = (uint32_t)ov_quads.size();
      ov_cells[start + cl.first * 2 + 1] = (uint32_t)cl.second.size();
      ov_quads.insert(ov_quads.end(), cl.second.begin(), cl.second.end());
    }
  }
  out.overlay_bricks = ov_cells.size() / OV_CELL_WORDS;
  out.see_through[0] = entity_base + (uint32_t)v.entity_quads_len - v.see_through_count;
  out.see_through[1] = v.see_through_count;

  emitter_list(v.models, sites, out.emitters, out.emitter_grid);
  out.emitter_count = (uint32_t)sites.size();

  // :427-440
  std::vector<uint32_t> &g = out.grid;
  out.offsets[0] = 0;
  g.insert(g.end(), top.begin(), top.end());
  out.offsets[1] = (uint32_t)g.size();
  g.insert(g.end(), bricks.begin(), bricks.end());
  out.offsets[2] = (uint32_t)g.size();
  g.insert(g.end(), ov_top.begin(), ov_top.end());
  out.offsets[3] = (uint32_t)g.size();
  g.insert(g.end(), ov_cells.begin(), ov_cells.end());
  out.offsets[4] = (uint32_t)g.size();
  g.insert(g.end(), ov_quads.begin(), ov_quads.end());
  g.push_back(0u);
  return out;
}

}  // namespace lwo_build
