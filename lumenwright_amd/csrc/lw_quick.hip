// Quick mode on the device: the reference's tile-binned rasterizer
// (crates/lumenwright/src/render/cpu/raster.rs) with the client's terrain sampling
// (crates/lumenwright/src/render/texture.rs), in the same f32 operations and order, so a frame
// is the CPU rasterizer's pixel for pixel.
//
//   k_verts        one lane per vertex: the view-projection and the attributes (raster.rs:176-193)
//   k_tris         one lane per triangle of the three passes: trivial reject, near-plane
//                  clipping into up to two triangles, screen set-up (raster.rs:51-135,195-227),
//                  into two slots per triangle, so slot order is the reference's triangle order
//   k_bin_counts,  one lane per slot: a (tile, slot) key per 8 x 8 tile its box touches, radix-
//   k_bin_keys     sorted, so a tile's list is in triangle order (raster.rs:234-252 bins by
//                  32 x 32 tiles; a pixel sees the same triangles in the same order either way)
//   k_tile_bounds  each tile's run of the sorted keys (a binary search a tile)
//   k_tiles        one wave per 8 x 8 tile, one pixel a lane: the sky, then the tile's list in
//                  order. Each lane queues the triangles that cover its pixel (edge and depth-
//                  range tests) in LDS, and the wave shades the queues in lockstep: the depth
//                  test, the atlas sampling, the fog and the blending of raster.rs:266-357, in
//                  each pixel's triangle order.
#include "lw_quick.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <climits>

#include "lw_libm.hpp"
#include "lw_math.hpp"

namespace lwq {
namespace {

constexpr uint32_t kB = 256;
constexpr uint32_t kTile = 8;    // device bins: one wave64 a tile, one pixel a lane
constexpr uint32_t kWaves = kB / 64;
constexpr uint32_t kQueue = 8;   // covering triangles a lane holds before its wave shades them
constexpr float kMaxAnisotropy = 4.0f;  // assets/atlas.rs:25-27

#define LWQ_TRY(expr)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) throw Error{false, e_, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
  } while (0)

template <class T>
struct Buf {
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    free();
    if (count) {
      LWQ_TRY(hipMalloc((void **)&p, count * sizeof(T)));
      n = count;
    }
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~Buf() { free(); }
};

using lw::V3;
using lw::v3;

struct ClipVert {
  float4 clip;
  float uv[2];
  float4 color;
  float pos[3];
};

// raster.rs:24-36
struct Tri {
  float4 v[3];     // screen x, y, ndc z, 1 / clip w
  float a[3][9];   // uv, colour, position, each divided by clip w
  float area_inv;
  uint32_t flags;  // 1 translucent, 2 see-through
  int32_t bbox[4];
};
constexpr uint32_t TRI_TRANSLUCENT = 1u, TRI_SEE_THROUGH = 2u;

__device__ __forceinline__ float4 f4add(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ __forceinline__ float4 f4sub(float4 a, float4 b) { return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
__device__ __forceinline__ float4 f4mul(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }
__device__ __forceinline__ float4 f4div(float4 a, float s) { return make_float4(a.x / s, a.y / s, a.z / s, a.w / s); }
// glam 0.34.1's lerp: a * (1 - s) + b * s
__device__ __forceinline__ float4 f4lerp(float4 a, float4 b, float s) { return f4add(f4mul(a, 1.0f - s), f4mul(b, s)); }
__device__ __forceinline__ float rmin(float a, float b) { return isnan(a) ? b : (isnan(b) ? a : (a < b ? a : b)); }
__device__ __forceinline__ float rmax(float a, float b) { return isnan(a) ? b : (isnan(b) ? a : (a > b ? a : b)); }
__device__ __forceinline__ float rclamp(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }  // f32::clamp

// raster.rs:39-50
__device__ __forceinline__ ClipVert lerp_vert(const ClipVert &a, const ClipVert &b, float t) {
  ClipVert r;
  r.clip = f4add(a.clip, f4mul(f4sub(b.clip, a.clip), t));
  r.uv[0] = a.uv[0] + (b.uv[0] - a.uv[0]) * t;
  r.uv[1] = a.uv[1] + (b.uv[1] - a.uv[1]) * t;
  r.color = f4add(a.color, f4mul(f4sub(b.color, a.color), t));
  for (int k = 0; k < 3; k++) r.pos[k] = a.pos[k] + (b.pos[k] - a.pos[k]) * t;
  return r;
}

// raster.rs:176-193: clip = view_proj * (pos, 1), glam's column order; colour / 255.
__global__ void __launch_bounds__(kB) k_verts(const lw_quick_vertex *__restrict__ in, uint32_t n,
                                              const Input q, ClipVert *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const lw_quick_vertex v = in[i];
  const float *m = q.view_proj;
  float4 c = make_float4(m[0] * v.pos[0], m[1] * v.pos[0], m[2] * v.pos[0], m[3] * v.pos[0]);
  c = f4add(c, make_float4(m[4] * v.pos[1], m[5] * v.pos[1], m[6] * v.pos[1], m[7] * v.pos[1]));
  c = f4add(c, make_float4(m[8] * v.pos[2], m[9] * v.pos[2], m[10] * v.pos[2], m[11] * v.pos[2]));
  c = f4add(c, make_float4(m[12] * 1.0f, m[13] * 1.0f, m[14] * 1.0f, m[15] * 1.0f));
  ClipVert r;
  r.clip = c;
  r.uv[0] = v.uv[0];
  r.uv[1] = v.uv[1];
  r.color = f4div(make_float4((float)v.color[0], (float)v.color[1], (float)v.color[2], (float)v.color[3]), 255.0f);
  for (int k = 0; k < 3; k++) r.pos[k] = v.pos[k];
  out[i] = r;
}

// raster.rs:83-135. Returns false for a culled or empty triangle.
__device__ bool setup(const ClipVert *v, float w, float h, uint32_t flags, Tri &t) {
  float4 sv[3];
  float attrs[3][9];
  for (int i = 0; i < 3; i++) {
    const float4 c = v[i].clip;
    if (c.w <= 1e-6f) return false;
    const float iw = 1.0f / c.w;
    const float nx = c.x * iw, ny = c.y * iw, nz = c.z * iw;
    sv[i] = make_float4((nx * 0.5f + 0.5f) * w, (0.5f - ny * 0.5f) * h, nz, iw);
    const float src[9] = {v[i].uv[0], v[i].uv[1], v[i].color.x, v[i].color.y, v[i].color.z,
                          v[i].color.w, v[i].pos[0], v[i].pos[1], v[i].pos[2]};
    for (int k = 0; k < 9; k++) attrs[i][k] = src[k] * iw;
  }
  // Rows grow downward on screen: a front-facing (counter-clockwise) triangle comes out with a
  // negative signed area here, and anything else is culled.
  const float area = (sv[1].x - sv[0].x) * (sv[2].y - sv[0].y) - (sv[1].y - sv[0].y) * (sv[2].x - sv[0].x);
  if (area >= 0.0f) return false;
  float min_x = 3.40282347e38f, max_x = -3.40282347e38f, min_y = 3.40282347e38f, max_y = -3.40282347e38f;
  for (int i = 0; i < 3; i++) {
    min_x = rmin(min_x, sv[i].x);
    max_x = rmax(max_x, sv[i].x);
    min_y = rmin(min_y, sv[i].y);
    max_y = rmax(max_y, sv[i].y);
  }
  const int32_t wi = lw::as_i32(w), hi = lw::as_i32(h);
  t.bbox[0] = max(lw::as_i32(floorf(min_x)), 0);
  t.bbox[1] = max(lw::as_i32(floorf(min_y)), 0);
  t.bbox[2] = min(lw::as_i32(ceilf(max_x)), wi - 1);
  t.bbox[3] = min(lw::as_i32(ceilf(max_y)), hi - 1);
  if (t.bbox[0] > t.bbox[2] || t.bbox[1] > t.bbox[3]) return false;
  // Swapping the last two vertices makes the signed area positive, as the edge functions
  // below expect.
  const int order[3] = {0, 2, 1};
  for (int i = 0; i < 3; i++) {
    t.v[i] = sv[order[i]];
    for (int k = 0; k < 9; k++) t.a[i][k] = attrs[order[i]][k];
  }
  t.area_inv = 1.0f / -area;
  t.flags = flags;
  return true;
}

// raster.rs:195-227 and clip_near (raster.rs:52-81): triangle g of the concatenated lists
// into slots 2g and 2g + 1 (valid[] says which hold one).
__global__ void __launch_bounds__(kB) k_tris(const uint32_t *__restrict__ idx, uint32_t n_tri, uint32_t n_opaque,
                                             uint32_t n_translucent, const ClipVert *__restrict__ verts,
                                             float w, float h, Tri *__restrict__ slots,
                                             uint32_t *__restrict__ valid) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_tri) return;
  valid[2 * g] = 0u;
  valid[2 * g + 1] = 0u;
  const uint32_t flags = g < n_opaque ? 0u
                         : (g < n_opaque + n_translucent ? TRI_TRANSLUCENT : (TRI_TRANSLUCENT | TRI_SEE_THROUGH));
  ClipVert tv[3];
  for (int k = 0; k < 3; k++) tv[k] = verts[idx[3 * (size_t)g + k]];
  // A triangle whose three vertices all lie beyond the same clip plane is dropped at once.
  bool out[5] = {true, true, true, true, true};
  for (int k = 0; k < 3; k++) {
    const float4 c = tv[k].clip;
    out[0] = out[0] && c.x > c.w;
    out[1] = out[1] && c.x < -c.w;
    out[2] = out[2] && c.y > c.w;
    out[3] = out[3] && c.y < -c.w;
    out[4] = out[4] && c.z < 0.0f;
  }
  if (out[0] || out[1] || out[2] || out[3] || out[4]) return;
  ClipVert poly[4];
  int np = 0, inside_n = 0;
  for (int k = 0; k < 3; k++) inside_n += tv[k].clip.z >= 0.0f ? 1 : 0;
  if (inside_n == 3) {
    poly[0] = tv[0];
    poly[1] = tv[1];
    poly[2] = tv[2];
    np = 3;
  } else if (inside_n > 0) {
    for (int i = 0; i < 3; i++) {
      const ClipVert &a = tv[i], &b = tv[(i + 1) % 3];
      const bool ia = a.clip.z >= 0.0f, ib = b.clip.z >= 0.0f;
      if (ia) poly[np++] = a;
      if (ia != ib) {
        const float t = a.clip.z / (a.clip.z - b.clip.z);
        poly[np++] = lerp_vert(a, b, t);
      }
    }
  }
  for (int i = 1; i + 1 < np; i++) {
    const ClipVert c[3] = {poly[0], poly[i], poly[i + 1]};
    Tri t;
    if (setup(c, w, h, flags, t)) {
      slots[2 * (size_t)g + (i - 1)] = t;
      valid[2 * g + (i - 1)] = 1u;
    }
  }
}

// The tiles a slot's box touches (0 for an empty slot).
__global__ void __launch_bounds__(kB) k_bin_counts(const Tri *__restrict__ slots, const uint32_t *__restrict__ valid,
                                                   size_t n, unsigned long long *__restrict__ counts) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    uint64_t c = 0;
    if (valid[i]) {
      const int32_t *b = slots[i].bbox;
      const uint64_t nx = (uint64_t)((uint32_t)b[2] / kTile - (uint32_t)b[0] / kTile + 1);
      const uint64_t ny = (uint64_t)((uint32_t)b[3] / kTile - (uint32_t)b[1] / kTile + 1);
      c = nx * ny;
    }
    counts[i] = c;
  }
}

__global__ void __launch_bounds__(kB) k_bin_keys(const Tri *__restrict__ slots, const uint32_t *__restrict__ valid,
                                                 size_t n, const unsigned long long *__restrict__ off,
                                                 uint32_t tiles_x, unsigned long long *__restrict__ keys) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if (!valid[i]) continue;
    const int32_t *b = slots[i].bbox;
    unsigned long long j = off[i];
    for (uint32_t ty = (uint32_t)b[1] / kTile; ty <= (uint32_t)b[3] / kTile; ty++)
      for (uint32_t tx = (uint32_t)b[0] / kTile; tx <= (uint32_t)b[2] / kTile; tx++)
        keys[j++] = ((unsigned long long)(ty * tiles_x + tx) << 32) | (unsigned long long)i;
  }
}

// start[t] = the first sorted key of tile t (n_keys for t = n_tiles): a binary search a tile.
__global__ void __launch_bounds__(kB) k_tile_bounds(const unsigned long long *__restrict__ sorted, uint32_t n_keys,
                                                    uint32_t n_tiles, uint32_t *__restrict__ start) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n_tiles) return;
  uint32_t lo = 0, hi = n_keys;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if ((uint32_t)(sorted[mid] >> 32) < t) lo = mid + 1; else hi = mid;
  }
  start[t] = lo;
}

__global__ void __launch_bounds__(kB) k_fill_u32(uint32_t *__restrict__ p, size_t n, uint32_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---- the atlas (texture.rs) ----

struct Atlas {
  const uint32_t *level0, *mips;
  uint32_t w, h, levels, mip_rows;
  const uint32_t *mip_off;
  const uint2 *log2_fix;
  uint32_t n_log2_fix;
};

__device__ __forceinline__ float4 unpack(uint32_t c) {
  return f4div(make_float4((float)(c & 0xFFu), (float)((c >> 8) & 0xFFu), (float)((c >> 16) & 0xFFu),
                           (float)(c >> 24)),
               255.0f);
}

// raster.rs:163-167: clamp (max then min), * 255 + 0.5, saturating to u32.
__device__ __forceinline__ uint32_t pack(float r, float g, float b) {
  const float c[3] = {r, g, b};
  uint32_t o[3];
  for (int k = 0; k < 3; k++) o[k] = lw::as_u32(rmin(rmax(c[k], 0.0f), 1.0f) * 255.0f + 0.5f);
  return o[0] | (o[1] << 8) | (o[2] << 16) | 0xFF000000u;
}

// texture.rs:48-54
__device__ __forceinline__ float4 nearest(const Atlas &A, float u, float v) {
  const int32_t x = min(max(lw::as_i32(floorf(u * (float)A.w)), 0), (int32_t)A.w - 1);
  const int32_t y = min(max(lw::as_i32(floorf(v * (float)A.h)), 0), (int32_t)A.h - 1);
  return unpack(A.level0[(uint32_t)y * A.w + (uint32_t)x]);
}

// texture.rs:57-59
__device__ __forceinline__ bool is_mipmapped(const Atlas &A, float v) {
  return A.levels > 0 && v * (float)A.h < (float)A.mip_rows;
}

// texture.rs:61-73
__device__ __forceinline__ float4 texel(const Atlas &A, uint32_t l, int32_t x, int32_t y) {
  const int32_t w = (int32_t)(A.w >> l), h = (int32_t)(A.h >> l);
  const uint32_t i = (uint32_t)(min(max(y, 0), h - 1) * w + min(max(x, 0), w - 1));
  return unpack(l == 0 ? A.level0[i] : A.mips[A.mip_off[l] + i]);
}

// texture.rs:75-86
__device__ __forceinline__ float4 bilinear(const Atlas &A, float u, float v, uint32_t l) {
  const float sx = (float)(A.w >> l), sy = (float)(A.h >> l);
  const float tx = u * sx - 0.5f, ty = v * sy - 0.5f;
  const float fx = floorf(tx), fy = floorf(ty);
  const float rx = tx - fx, ry = ty - fy;
  const int32_t x = lw::as_i32(fx), y = lw::as_i32(fy);
  // x + 1 and y + 1 wrap at i32::MAX, as the reference's release build does.
  const int32_t x1 = (int32_t)((uint32_t)x + 1u), y1 = (int32_t)((uint32_t)y + 1u);
  const float4 top = f4lerp(texel(A, l, x, y), texel(A, l, x1, y), rx);
  const float4 bottom = f4lerp(texel(A, l, x, y1), texel(A, l, x1, y1), rx);
  return f4lerp(top, bottom, ry);
}

// texture.rs:88-98
__device__ __forceinline__ float4 trilinear(const Atlas &A, float u, float v, float lod) {
  const float lo = floorf(lod);
  const float fr = lod - lo;
  const uint32_t l = lw::as_u32(lo);
  const float4 a = bilinear(A, u, v, l);
  if (fr > 0.0f && l < A.levels) return f4lerp(a, bilinear(A, u, v, l + 1), fr);
  return a;
}

// texture.rs:102-119
__device__ float4 sample_grad(const Atlas &A, float u, float v, float dux, float duy, float dvx, float dvy) {
  const float sx = (float)A.w, sy = (float)A.h;
  const float ax = dux * sx, ay = duy * sy, bx = dvx * sx, by = dvy * sy;
  const float px = sqrtf((ax * ax) + (ay * ay)), py = sqrtf((bx * bx) + (by * by));
  const float pmax = rmax(px, py), pmin = rmin(px, py);
  const float n = rclamp(ceilf(pmax / rmax(pmin, 1e-8f)), 1.0f, kMaxAnisotropy);
  const float lod = rclamp(lw::log2_libm(rmax(pmax / n, 1e-8f), A.log2_fix, A.n_log2_fix), 0.0f, (float)A.levels);
  const bool along_u = px >= py;
  const float axx = along_u ? dux : dvx, axy = along_u ? duy : dvy;
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const uint32_t taps = lw::as_u32(n);
  for (uint32_t i = 0; i < taps; i++) {
    const float off = ((float)i + 0.5f) / n - 0.5f;
    acc = f4add(acc, trilinear(A, u + axx * off, v + axy * off, lod));
  }
  return f4div(acc, n);
}

// texture.rs:121-133: the client's sampleNearest, then textureGrad.
__device__ float4 sample_block(const Atlas &A, float u, float v, float dux, float duy, float dvx, float dvy) {
  const float sx = (float)A.w, sy = (float)A.h;
  const float pxl = 1.0f / sx, pyl = 1.0f / sy;
  const float tx = u * sx, ty = v * sy;
  const float cx = rintf(tx) - 0.5f, cy = rintf(ty) - 0.5f;
  const float ox = tx - cx, oy = ty - cy;
  const float scx = sqrtf(rmax((dux * dux) + (dvx * dvx), 1e-20f));
  const float scy = sqrtf(rmax((duy * duy) + (dvy * dvy), 1e-20f));
  const float offx = rmin(rmax((ox - 0.5f) * pxl / scx + 0.5f, 0.0f), 1.0f);
  const float offy = rmin(rmax((oy - 0.5f) * pyl / scy + 0.5f, 0.0f), 1.0f);
  return sample_grad(A, (cx + offx) * pxl, (cy + offy) * pyl, dux, duy, dvx, dvy);
}

__device__ __forceinline__ float edge(float4 a, float4 b, float px, float py) {
  return (b.x - a.x) * (py - a.y) - (b.y - a.y) * (px - a.x);
}

// QuickSky::pixel_dir (sky.rs:376-384)
__device__ __forceinline__ V3 pixel_dir(const lw_quick_sky &q, float x, float y) {
  const float *m = q.cam_misc;
  const float sx = (2.0f * (x + 0.5f) / m[2] - 1.0f) * m[1] * m[0];
  const float sy = (1.0f - 2.0f * (y + 0.5f) / m[3]) * m[0];
  const V3 f = lw::xyz(q.cam_forward), r = lw::xyz(q.cam_right), u = lw::xyz(q.cam_up);
  const V3 d = f + r * sx + u * sy;
  const float inv = 1.0f / sqrtf(lw::dot(d, d));
  return d * inv;
}

// QuickSky::fog_factor (sky.rs:537-545)
__device__ __forceinline__ float fog_factor(const lw_quick_sky &q, float x, float y, float z) {
  const float sph = sqrtf((x * x) + (y * y) + (z * z));
  const float cyl = rmax(sqrtf(x * x + z * z), fabsf(y));
  return rmax(lw::linear_fog(sph, q.fog_dist[1], q.fog_dist[2]), lw::linear_fog(cyl, q.fog_color[3], q.fog_dist[0]));
}

// Whether triangle t covers pixel (px, py) by the tests that do not depend on earlier
// fragments: the box, the three edges and the depth range (raster.rs:285-306; a NaN depth
// passes, as there).
__device__ __forceinline__ bool covers(const Tri &t, uint32_t px, uint32_t py) {
  if ((int32_t)px < t.bbox[0] || (int32_t)px > t.bbox[2] || (int32_t)py < t.bbox[1] || (int32_t)py > t.bbox[3])
    return false;
  const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
  const float w0 = edge(t.v[1], t.v[2], fx, fy);
  const float w1 = edge(t.v[2], t.v[0], fx, fy);
  const float w2 = edge(t.v[0], t.v[1], fx, fy);
  if (w0 < 0.0f || w1 < 0.0f || w2 < 0.0f) return false;
  const float z = w0 * t.area_inv * t.v[0].z + w1 * t.area_inv * t.v[1].z + w2 * t.area_inv * t.v[2].z;
  return !(z < 0.0f || z > 1.0f);
}

// One covering fragment of triangle t at pixel (px, py) (raster.rs:293-354): the depth test,
// the perspective-correct attributes, the atlas sample, the fog and the blend or the write.
__device__ void shade(const Tri &t, uint32_t px, uint32_t py, const Atlas &A, const lw_quick_sky &sky,
                      const V3 fog, float &depth, uint32_t &color) {
  const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
  const float w0 = edge(t.v[1], t.v[2], fx, fy);
  const float w1 = edge(t.v[2], t.v[0], fx, fy);
  const float w2 = edge(t.v[0], t.v[1], fx, fy);
  const float b0 = w0 * t.area_inv, b1 = w1 * t.area_inv, b2 = w2 * t.area_inv;
  const float z = b0 * t.v[0].z + b1 * t.v[1].z + b2 * t.v[2].z;
  const bool see_through = (t.flags & TRI_SEE_THROUGH) != 0u, translucent = (t.flags & TRI_TRANSLUCENT) != 0u;
  if (z >= depth && !see_through) return;
  const float iw = b0 * t.v[0].w + b1 * t.v[1].w + b2 * t.v[2].w;
  const float pw = 1.0f / iw;
  float at[9];
  for (int a = 0; a < 9; a++) at[a] = (b0 * t.a[0][a] + b1 * t.a[1][a] + b2 * t.a[2][a]) * pw;
  // The UVs the triangle's plane gives at a neighbouring pixel centre, as a quad of GPU
  // fragments would see them in its helper lanes.
  auto uv_at = [&](float x, float y, float &u, float &v) {
    const float c0 = edge(t.v[1], t.v[2], x, y) * t.area_inv;
    const float c1 = edge(t.v[2], t.v[0], x, y) * t.area_inv;
    const float c2 = edge(t.v[0], t.v[1], x, y) * t.area_inv;
    const float w = 1.0f / (c0 * t.v[0].w + c1 * t.v[1].w + c2 * t.v[2].w);
    u = (c0 * t.a[0][0] + c1 * t.a[1][0] + c2 * t.a[2][0]) * w;
    v = (c0 * t.a[0][1] + c1 * t.a[1][1] + c2 * t.a[2][1]) * w;
  };
  float4 tex;
  if (is_mipmapped(A, at[1])) {
    // Fine derivatives: the difference across this pixel's 2x2 quad.
    const float qx = (float)(px & ~1u) + 0.5f, qy = (float)(py & ~1u) + 0.5f;
    float u1, v1, u0, v0, u2, v2, u3, v3;
    uv_at(qx + 1.0f, fy, u1, v1);
    uv_at(qx, fy, u0, v0);
    uv_at(fx, qy + 1.0f, u2, v2);
    uv_at(fx, qy, u3, v3);
    tex = sample_block(A, at[0], at[1], u1 - u0, v1 - v0, u2 - u3, v2 - v3);
  } else {
    tex = nearest(A, at[0], at[1]);
  }
  V3 rgb = v3(tex.x * at[2], tex.y * at[3], tex.z * at[4]);
  rgb = lw::lerp3(rgb, fog, fog_factor(sky, at[6], at[7], at[8]));
  if (translucent) {
    const float a = tex.w * at[5];
    if (a <= 0.0f) return;
    const float4 d = unpack(color);
    const V3 mix = lw::lerp3(v3(d.x, d.y, d.z), rgb, a);
    color = pack(mix.x, mix.y, mix.z);
  } else {
    if (tex.w < 0.5f) return;
    depth = z;
    color = pack(rgb.x, rgb.y, rgb.z);
  }
}

// raster.rs:266-372 for one 8 x 8 tile per wave: the sky, then the tile's triangles in order.
// A lane queues the triangles that cover its pixel; when a queue is full, and at the end, the
// wave shades every lane's queue in order, so each lane shades its own fragment in lockstep
// with the others rather than waiting on the few lanes a small triangle covers.
__global__ void __launch_bounds__(kB) k_tiles(const Tri *__restrict__ slots, const unsigned long long *__restrict__ keys,
                                              const uint32_t *__restrict__ tile_start, const Input q,
                                              uint32_t tiles_x, uint32_t n_tiles, uint32_t *__restrict__ out) {
  __shared__ uint32_t queue[kWaves][kQueue][64];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t tile = blockIdx.x * kWaves + wave;
  if (tile >= n_tiles) return;  // the whole wave
  const uint32_t px = (tile % tiles_x) * kTile + lane % kTile, py = (tile / tiles_x) * kTile + lane / kTile;
  const bool live = px < q.width && py < q.height;
  const Atlas A{q.atlas, q.mips, q.atlas_w, q.atlas_h, q.levels, q.mip_rows, q.mip_offset, q.log2_fix,
                q.n_log2_fix};
  const V3 fog = lw::xyz(q.sky.fog_color);
  uint32_t color = 0u;
  float depth = INFINITY;
  if (live) {
    const V3 c = lw::qs_background(q.sky, q.sectors, pixel_dir(q.sky, (float)px, (float)py), q.atlas, q.atlas_w,
                                   q.atlas_w * q.atlas_h, q.stars);
    color = pack(c.x, c.y, c.z);
  }
  uint32_t e = tile_start[tile];
  const uint32_t end = tile_start[tile + 1];
  uint32_t qlen = 0;
  for (;;) {
    while (e < end) {
      const uint32_t s = __builtin_amdgcn_readfirstlane((uint32_t)keys[e]);
      e++;
      if (live && covers(slots[s], px, py)) queue[wave][qlen++][lane] = s;
      if (__ballot(qlen == kQueue) != 0ull) break;
    }
    for (uint32_t j = 0; __ballot(j < qlen) != 0ull; j++)
      if (j < qlen) shade(slots[queue[wave][j][lane]], px, py, A, q.sky, fog, depth, color);
    qlen = 0;
    if (e >= end) break;
  }
  if (live) out[(size_t)py * q.width + px] = color;
}

unsigned grid_for(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + kB - 1) / kB, 1u << 20)); }

uint32_t bit_width(uint64_t v) {
  uint32_t b = 0;
  while (v) {
    b++;
    v >>= 1;
  }
  return b;
}

}  // namespace

// Grow-only device buffers kept between calls.
struct Workspace {
  Buf<ClipVert> verts;
  Buf<Tri> slots;
  Buf<uint32_t> valid, tile_start;
  Buf<unsigned long long> counts, off, keys, sorted;
  Buf<char> temp;
};

Workspace *workspace_new() { return new Workspace(); }
void workspace_free(Workspace *ws) { delete ws; }

namespace {
template <class T>
T *grow(Buf<T> &b, size_t n) {
  if (n > b.n) b.alloc(n);
  return b.p;
}
}  // namespace

void render(hipStream_t s, const Input &in, Workspace &ws, uint32_t *out, Stats &st) {
  st = Stats{};
  const uint64_t n_tri = (uint64_t)in.n_opaque + in.n_translucent + in.n_see_through;
  if (n_tri >= (1ull << 30)) throw Error{true, hipSuccess, "too many triangles"};
  const uint32_t tiles_x = (in.width + kTile - 1) / kTile, tiles_y = (in.height + kTile - 1) / kTile;
  const uint64_t n_tiles = (uint64_t)tiles_x * tiles_y;
  const size_t n_slots = 2 * (size_t)n_tri;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  struct Events {
    hipEvent_t *a, *b;
    ~Events() {
      if (*a) (void)hipEventDestroy(*a);
      if (*b) (void)hipEventDestroy(*b);
    }
  } events{&ev0, &ev1};
  LWQ_TRY(hipEventCreate(&ev0));
  LWQ_TRY(hipEventCreate(&ev1));
  // Every buffer up to the sorted keys' size is known before the first launch.
  ClipVert *verts = grow(ws.verts, std::max<size_t>(in.n_verts, 1));
  Tri *slots = grow(ws.slots, std::max<size_t>(n_slots, 1));
  uint32_t *valid = grow(ws.valid, std::max<size_t>(n_slots, 1));
  unsigned long long *counts = grow(ws.counts, std::max<size_t>(n_slots, 1));
  unsigned long long *off = grow(ws.off, std::max<size_t>(n_slots, 1));
  uint32_t *tile_start = grow(ws.tile_start, n_tiles + 1);
  size_t scan_bytes = 0;
  LWQ_TRY(rocprim::exclusive_scan(nullptr, scan_bytes, counts, off, 0ull, std::max<size_t>(n_slots, 1),
                                  rocprim::plus<unsigned long long>(), s));
  grow(ws.temp, std::max<size_t>(scan_bytes, 1));
  LWQ_TRY(hipEventRecord(ev0, s));
  if (in.n_verts) {
    k_verts<<<dim3((in.n_verts + kB - 1) / kB), dim3(kB), 0, s>>>(in.verts, in.n_verts, in, verts);
    LWQ_TRY(hipGetLastError());
  }
  uint64_t n_keys = 0;
  if (n_tri) {
    // Triangles into two slots each (slot order is triangle order), their tile counts and
    // the keys' offsets; the key count is the one value read back before the last launches.
    k_tris<<<dim3((unsigned)((n_tri + kB - 1) / kB)), dim3(kB), 0, s>>>(
        in.indices, (uint32_t)n_tri, in.n_opaque, in.n_translucent, verts, (float)in.width, (float)in.height,
        slots, valid);
    LWQ_TRY(hipGetLastError());
    k_bin_counts<<<dim3(grid_for(n_slots)), dim3(kB), 0, s>>>(slots, valid, n_slots, counts);
    LWQ_TRY(hipGetLastError());
    LWQ_TRY(rocprim::exclusive_scan(ws.temp.p, scan_bytes, counts, off, 0ull, n_slots,
                                    rocprim::plus<unsigned long long>(), s));
    unsigned long long tail[2];
    LWQ_TRY(hipMemcpyAsync(&tail[0], off + n_slots - 1, sizeof(tail[0]), hipMemcpyDeviceToHost, s));
    LWQ_TRY(hipMemcpyAsync(&tail[1], counts + n_slots - 1, sizeof(tail[1]), hipMemcpyDeviceToHost, s));
    LWQ_TRY(hipStreamSynchronize(s));
    n_keys = tail[0] + tail[1];
  }
  if (n_keys >= (1ull << 32)) throw Error{true, hipSuccess, "triangles cover too many tiles"};
  st.bin_entries = n_keys;
  if (n_keys) {
    unsigned long long *keys = grow(ws.keys, n_keys), *sorted = grow(ws.sorted, n_keys);
    k_bin_keys<<<dim3(grid_for(n_slots)), dim3(kB), 0, s>>>(slots, valid, n_slots, off, tiles_x, keys);
    LWQ_TRY(hipGetLastError());
    // The keys are written in slot order, and the radix sort is stable: sorting the tile bits
    // alone keeps each tile's slots in order.
    size_t bytes = 0;
    const uint32_t end_bit = 32 + std::max(1u, bit_width(n_tiles - 1));
    LWQ_TRY(rocprim::radix_sort_keys(nullptr, bytes, keys, sorted, n_keys, 32, end_bit, s));
    LWQ_TRY(rocprim::radix_sort_keys(grow(ws.temp, std::max<size_t>(bytes, 1)), bytes, keys, sorted, n_keys, 32,
                                     end_bit, s));
    k_tile_bounds<<<dim3((unsigned)((n_tiles + kB) / kB)), dim3(kB), 0, s>>>(sorted, (uint32_t)n_keys,
                                                                             (uint32_t)n_tiles, tile_start);
  } else {
    k_fill_u32<<<dim3(grid_for(n_tiles + 1)), dim3(kB), 0, s>>>(tile_start, n_tiles + 1, 0u);
  }
  LWQ_TRY(hipGetLastError());
  k_tiles<<<dim3((unsigned)((n_tiles + kWaves - 1) / kWaves)), dim3(kB), 0, s>>>(
      slots, ws.sorted.p, tile_start, in, tiles_x, (uint32_t)n_tiles, out);
  LWQ_TRY(hipGetLastError());
  LWQ_TRY(hipEventRecord(ev1, s));
  LWQ_TRY(hipStreamSynchronize(s));
  float ms = 0.0f;
  LWQ_TRY(hipEventElapsedTime(&ms, ev0, ev1));
  st.kernel_ms = ms;
}

}  // namespace lwq
