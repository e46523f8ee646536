// ORACLE -- test infrastructure only. A restatement of the reference's quick-mode CPU
// rasterizer, crates/lumenwright/src/render/cpu/raster.rs (near-plane clipping :52-81,
// triangle set-up :83-135, the three passes, 32 x 32 tile bins and the per-pixel depth test,
// sampling, fog and blending :168-372), and of its atlas sampling,
// crates/lumenwright/src/render/texture.rs (nearest, bilinear, trilinear, the anisotropic
// textureGrad and the client's sampleNearest :46-145). f32 with libm's log2f; glam 0.34.1's
// Vec4 and Mat4 arithmetic (lerp a * (1 - s) + b * s; Mat4 * v = c0 v.x + c1 v.y + c2 v.z +
// c3 v.w). The sky behind the geometry is lw_oracle_sky.hpp's QuickSky::background.
#pragma once

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "lw_oracle_sky.hpp"

namespace lwo_quick {

using lwo_gsky::Quick;

struct Vertex {
  float pos[3], uv[2];
  uint8_t color[4];
};
static_assert(sizeof(Vertex) == 24, "Vertex layout");

struct V4 {
  float x, y, z, w;
  V4 operator+(V4 b) const { return {x + b.x, y + b.y, z + b.z, w + b.w}; }
  V4 operator-(V4 b) const { return {x - b.x, y - b.y, z - b.z, w - b.w}; }
  V4 operator*(float s) const { return {x * s, y * s, z * s, w * s}; }
  V4 operator/(float s) const { return {x / s, y / s, z / s, w / s}; }
};
static inline V4 lerp4(V4 a, V4 b, float s) { return a * (1.0f - s) + b * s; }
static inline float rmin(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a < b ? a : b)); }
static inline float rmax(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a > b ? a : b)); }
static inline float rclamp(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
static inline int32_t as_i32(float x) {
  if (std::isnan(x)) return 0;
  if (x >= 2147483648.0f) return INT32_MAX;
  if (x <= -2147483648.0f) return INT32_MIN;
  return (int32_t)x;
}
static inline uint32_t as_u32(float x) {
  if (!(x > 0.0f)) return 0;
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

struct ClipVert {
  V4 clip;
  float uv[2];
  V4 color;
  float pos[3];
};

struct Tri {
  V4 v[3];
  float a[3][9];
  float area_inv;
  bool translucent, see_through;
  int32_t bbox[4];
};

// raster.rs:39-50
static ClipVert lerp_vert(const ClipVert &a, const ClipVert &b, float t) {
  ClipVert r;
  r.clip = a.clip + (b.clip - a.clip) * t;
  for (int k = 0; k < 2; k++) r.uv[k] = a.uv[k] + (b.uv[k] - a.uv[k]) * t;
  r.color = a.color + (b.color - a.color) * t;
  for (int k = 0; k < 3; k++) r.pos[k] = a.pos[k] + (b.pos[k] - a.pos[k]) * t;
  return r;
}

// raster.rs:52-81
static void clip_near(const ClipVert v[3], std::vector<std::array<ClipVert, 3>> &out) {
  auto inside = [](const ClipVert &c) { return c.clip.z >= 0.0f; };
  int n = 0;
  for (int i = 0; i < 3; i++) n += inside(v[i]) ? 1 : 0;
  if (n == 3) {
    out.push_back({v[0], v[1], v[2]});
    return;
  }
  if (n == 0) return;
  std::vector<ClipVert> poly;
  for (int i = 0; i < 3; i++) {
    const ClipVert &a = v[i], &b = v[(i + 1) % 3];
    if (inside(a)) poly.push_back(a);
    if (inside(a) != inside(b)) poly.push_back(lerp_vert(a, b, a.clip.z / (a.clip.z - b.clip.z)));
  }
  for (size_t i = 1; i + 1 < poly.size(); i++) out.push_back({poly[0], poly[i], poly[i + 1]});
}

enum Pass { OPAQUE, TRANSLUCENT, SEE_THROUGH };

// raster.rs:83-135
static bool setup(const std::array<ClipVert, 3> &v, float w, float h, Pass pass, Tri &t) {
  V4 sv[3];
  float at[3][9];
  for (int i = 0; i < 3; i++) {
    const V4 c = v[i].clip;
    if (c.w <= 1e-6f) return false;
    const float iw = 1.0f / c.w;
    const float nx = c.x * iw, ny = c.y * iw, nz = c.z * iw;
    sv[i] = {(nx * 0.5f + 0.5f) * w, (0.5f - ny * 0.5f) * h, nz, iw};
    const float src[9] = {v[i].uv[0], v[i].uv[1], v[i].color.x, v[i].color.y, v[i].color.z, v[i].color.w,
                          v[i].pos[0], v[i].pos[1], v[i].pos[2]};
    for (int k = 0; k < 9; k++) at[i][k] = src[k] * iw;
  }
  const float area = (sv[1].x - sv[0].x) * (sv[2].y - sv[0].y) - (sv[1].y - sv[0].y) * (sv[2].x - sv[0].x);
  if (area >= 0.0f) return false;
  float mnx = 3.40282347e38f, mxx = -3.40282347e38f, mny = 3.40282347e38f, mxy = -3.40282347e38f;
  for (int i = 0; i < 3; i++) {
    mnx = rmin(mnx, sv[i].x);
    mxx = rmax(mxx, sv[i].x);
    mny = rmin(mny, sv[i].y);
    mxy = rmax(mxy, sv[i].y);
  }
  t.bbox[0] = std::max(as_i32(std::floor(mnx)), 0);
  t.bbox[1] = std::max(as_i32(std::floor(mny)), 0);
  t.bbox[2] = std::min(as_i32(std::ceil(mxx)), as_i32(w) - 1);
  t.bbox[3] = std::min(as_i32(std::ceil(mxy)), as_i32(h) - 1);
  if (t.bbox[0] > t.bbox[2] || t.bbox[1] > t.bbox[3]) return false;
  const int order[3] = {0, 2, 1};
  for (int i = 0; i < 3; i++) {
    t.v[i] = sv[order[i]];
    std::memcpy(t.a[i], at[order[i]], sizeof(t.a[i]));
  }
  t.area_inv = 1.0f / -area;
  t.translucent = pass != OPAQUE;
  t.see_through = pass == SEE_THROUGH;
  return true;
}

static inline float edge(V4 a, V4 b, float px, float py) { return (b.x - a.x) * (py - a.y) - (b.y - a.y) * (px - a.x); }

static inline V4 unpack(uint32_t c) {
  return V4{(float)(c & 0xFF), (float)((c >> 8) & 0xFF), (float)((c >> 16) & 0xFF), (float)(c >> 24)} / 255.0f;
}
static inline uint32_t pack(float r, float g, float b) {
  const float c[3] = {r, g, b};
  uint32_t o[3];
  for (int k = 0; k < 3; k++) o[k] = as_u32(rmin(rmax(c[k], 0.0f), 1.0f) * 255.0f + 0.5f);
  return o[0] | (o[1] << 8) | (o[2] << 16) | 0xFF000000u;
}

// texture.rs:21-145
struct AtlasView {
  const uint32_t *level0, *mips;
  uint32_t w, h, levels, mip_rows;
  V4 nearest(float u, float v) const {
    int32_t x = std::min(std::max(as_i32(std::floor(u * (float)w)), 0), (int32_t)w - 1);
    int32_t y = std::min(std::max(as_i32(std::floor(v * (float)h)), 0), (int32_t)h - 1);
    return unpack(level0[(uint32_t)y * w + (uint32_t)x]);
  }
  bool is_mipmapped(float v) const { return levels > 0 && v * (float)h < (float)mip_rows; }
  V4 texel(uint32_t l, int32_t x, int32_t y) const {
    int32_t lw = (int32_t)(w >> l), lh = (int32_t)(h >> l);
    size_t i = (size_t)(std::min(std::max(y, 0), lh - 1) * lw + std::min(std::max(x, 0), lw - 1));
    if (l == 0) return unpack(level0[i]);
    size_t o = 0;
    for (uint32_t k = 1; k < l; k++) o += (size_t)(w >> k) * (h >> k);
    return unpack(mips[o + i]);
  }
  V4 bilinear(float u, float v, uint32_t l) const {
    float sx = (float)(w >> l), sy = (float)(h >> l);
    float tx = u * sx - 0.5f, ty = v * sy - 0.5f;
    float fx = std::floor(tx), fy = std::floor(ty);
    float rx = tx - fx, ry = ty - fy;
    int32_t x = as_i32(fx), y = as_i32(fy);
    int32_t x1 = (int32_t)((uint32_t)x + 1u), y1 = (int32_t)((uint32_t)y + 1u);  // wrapping, as in release
    V4 top = lerp4(texel(l, x, y), texel(l, x1, y), rx);
    V4 bottom = lerp4(texel(l, x, y1), texel(l, x1, y1), rx);
    return lerp4(top, bottom, ry);
  }
  V4 trilinear(float u, float v, float lod) const {
    float lo = std::floor(lod), fr = lod - lo;
    uint32_t l = as_u32(lo);
    V4 a = bilinear(u, v, l);
    if (fr > 0.0f && l < levels) return lerp4(a, bilinear(u, v, l + 1), fr);
    return a;
  }
  V4 sample_grad(float u, float v, float dux, float duy, float dvx, float dvy) const {
    float sx = (float)w, sy = (float)h;
    float ax = dux * sx, ay = duy * sy, bx = dvx * sx, by = dvy * sy;
    float px = std::sqrt((ax * ax) + (ay * ay)), py = std::sqrt((bx * bx) + (by * by));
    float pmax = rmax(px, py), pmin = rmin(px, py);
    float n = rclamp(std::ceil(pmax / rmax(pmin, 1e-8f)), 1.0f, 4.0f);
    float lod = rclamp(std::log2(rmax(pmax / n, 1e-8f)), 0.0f, (float)levels);
    bool along_u = px >= py;
    float axx = along_u ? dux : dvx, axy = along_u ? duy : dvy;
    V4 acc{0, 0, 0, 0};
    uint32_t taps = as_u32(n);
    for (uint32_t i = 0; i < taps; i++) {
      float off = ((float)i + 0.5f) / n - 0.5f;
      acc = acc + trilinear(u + axx * off, v + axy * off, lod);
    }
    return acc / n;
  }
  V4 sample_block(float u, float v, float dux, float duy, float dvx, float dvy) const {
    float sx = (float)w, sy = (float)h;
    float pxl = 1.0f / sx, pyl = 1.0f / sy;
    float tx = u * sx, ty = v * sy;
    float cx = std::nearbyint(tx) - 0.5f, cy = std::nearbyint(ty) - 0.5f;
    float ox = tx - cx, oy = ty - cy;
    float scx = std::sqrt(rmax((dux * dux) + (dvx * dvx), 1e-20f));
    float scy = std::sqrt(rmax((duy * duy) + (dvy * dvy), 1e-20f));
    float offx = rmin(rmax((ox - 0.5f) * pxl / scx + 0.5f, 0.0f), 1.0f);
    float offy = rmin(rmax((oy - 0.5f) * pyl / scy + 0.5f, 0.0f), 1.0f);
    return sample_grad((cx + offx) * pxl, (cy + offy) * pyl, dux, duy, dvx, dvy);
  }
};

struct Frame {
  const Vertex *verts;
  size_t n_verts;
  const uint32_t *lists[3];
  size_t list_len[3];
  const uint32_t *atlas;
  uint32_t aw, ah;
  const uint32_t *mips;
  uint32_t levels, mip_rows;
  uint32_t width, height;
  float view_proj[16];
  const Quick *sky;
};

// QuickSky::pixel_dir (sky.rs:376-384) and fog_factor (sky.rs:537-545)
static lwo_gsky::V pixel_dir(const Quick &q, float x, float y) {
  const float *m = q.cam_misc;
  float sx = (2.0f * (x + 0.5f) / m[2] - 1.0f) * m[1] * m[0];
  float sy = (1.0f - 2.0f * (y + 0.5f) / m[3]) * m[0];
  return lwo_gsky::unit(lwo_gsky::v4(q.cam_forward) + lwo_gsky::v4(q.cam_right) * sx + lwo_gsky::v4(q.cam_up) * sy);
}
static float fog_factor(const Quick &q, float x, float y, float z) {
  float sph = std::sqrt((x * x) + (y * y) + (z * z));
  float cyl = rmax(std::sqrt(x * x + z * z), std::fabs(y));
  return rmax(lwo_gsky::linear_fog(sph, q.fog_dist[1], q.fog_dist[2]), lwo_gsky::linear_fog(cyl, q.fog_color[3], q.fog_dist[0]));
}

// raster.rs:168-372
static std::vector<uint32_t> rasterize(const Frame &f, const std::vector<uint32_t> &stars, unsigned threads = 1) {
  const float w = (float)f.width, h = (float)f.height;
  const float *m = f.view_proj;
  std::vector<ClipVert> verts(f.n_verts);
  for (size_t i = 0; i < f.n_verts; i++) {
    const Vertex &v = f.verts[i];
    V4 c = V4{m[0], m[1], m[2], m[3]} * v.pos[0];
    c = c + V4{m[4], m[5], m[6], m[7]} * v.pos[1];
    c = c + V4{m[8], m[9], m[10], m[11]} * v.pos[2];
    c = c + V4{m[12], m[13], m[14], m[15]} * 1.0f;
    verts[i].clip = c;
    verts[i].uv[0] = v.uv[0];
    verts[i].uv[1] = v.uv[1];
    verts[i].color = V4{(float)v.color[0], (float)v.color[1], (float)v.color[2], (float)v.color[3]} / 255.0f;
    for (int k = 0; k < 3; k++) verts[i].pos[k] = v.pos[k];
  }
  std::vector<Tri> tris;
  std::vector<std::array<ClipVert, 3>> clipped;
  for (int pass = 0; pass < 3; pass++)
    for (size_t t = 0; t + 3 <= f.list_len[pass]; t += 3) {  // a trailing partial triangle is dropped
      const uint32_t *ix = f.lists[pass] + t;
      const ClipVert tv[3] = {verts[ix[0]], verts[ix[1]], verts[ix[2]]};
      auto all = [&](auto pred) { return pred(tv[0].clip) && pred(tv[1].clip) && pred(tv[2].clip); };
      if (all([](V4 c) { return c.x > c.w; }) || all([](V4 c) { return c.x < -c.w; }) ||
          all([](V4 c) { return c.y > c.w; }) || all([](V4 c) { return c.y < -c.w; }) ||
          all([](V4 c) { return c.z < 0.0f; }))
        continue;
      clipped.clear();
      clip_near(tv, clipped);
      for (auto &c : clipped) {
        Tri tri;
        if (setup(c, w, h, (Pass)pass, tri)) tris.push_back(tri);
      }
    }
  const uint32_t TILE = 32;
  const uint32_t tiles_x = (f.width + TILE - 1) / TILE, tiles_y = (f.height + TILE - 1) / TILE;
  std::vector<std::vector<uint32_t>> bins((size_t)tiles_x * tiles_y);
  for (size_t i = 0; i < tris.size(); i++) {
    const Tri &t = tris[i];
    for (uint32_t ty = (uint32_t)t.bbox[1] / TILE; ty <= (uint32_t)t.bbox[3] / TILE; ty++)
      for (uint32_t tx = (uint32_t)t.bbox[0] / TILE; tx <= (uint32_t)t.bbox[2] / TILE; tx++)
        bins[(size_t)ty * tiles_x + tx].push_back((uint32_t)i);
  }
  const AtlasView view{f.atlas, f.mips, f.aw, f.ah, f.levels, f.mip_rows};
  const Quick &sky = *f.sky;
  const lwo_gsky::V fog = lwo_gsky::v4(sky.fog_color);
  std::vector<uint32_t> out((size_t)f.width * f.height);
  // The tiles in parallel, taken dynamically (raster.rs:263 maps them with rayon); each
  // writes only its own pixels.
  std::atomic<uint32_t> next_tile{0};
  auto worker = [&]() {
  for (uint32_t ti = next_tile.fetch_add(1); ti < tiles_x * tiles_y; ti = next_tile.fetch_add(1)) {
    const uint32_t x0 = (ti % tiles_x) * TILE, y0 = (ti / tiles_x) * TILE;
    const uint32_t tw = std::min(TILE, f.width - x0), th = std::min(TILE, f.height - y0);
    std::vector<uint32_t> color((size_t)tw * th);
    for (uint32_t py = 0; py < th; py++)
      for (uint32_t px = 0; px < tw; px++) {
        lwo_gsky::V c = lwo_gsky::background(sky, pixel_dir(sky, (float)(x0 + px), (float)(y0 + py)), f.atlas, f.aw, stars);
        color[(size_t)py * tw + px] = pack(c.x, c.y, c.z);
      }
    std::vector<float> depth((size_t)tw * th, INFINITY);
    for (uint32_t i : bins[ti]) {
      const Tri &t = tris[i];
      uint32_t bx0 = std::max((uint32_t)t.bbox[0], x0), by0 = std::max((uint32_t)t.bbox[1], y0);
      uint32_t bx1 = std::min((uint32_t)t.bbox[2], x0 + tw - 1), by1 = std::min((uint32_t)t.bbox[3], y0 + th - 1);
      if (bx0 > bx1 || by0 > by1) continue;
      for (uint32_t py = by0; py <= by1; py++)
        for (uint32_t px = bx0; px <= bx1; px++) {
          float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
          float w0 = edge(t.v[1], t.v[2], fx, fy), w1 = edge(t.v[2], t.v[0], fx, fy), w2 = edge(t.v[0], t.v[1], fx, fy);
          if (w0 < 0.0f || w1 < 0.0f || w2 < 0.0f) continue;
          float b0 = w0 * t.area_inv, b1 = w1 * t.area_inv, b2 = w2 * t.area_inv;
          float z = b0 * t.v[0].z + b1 * t.v[1].z + b2 * t.v[2].z;
          size_t li = (size_t)(py - y0) * tw + (px - x0);
          if ((z >= depth[li] && !t.see_through) || z < 0.0f || z > 1.0f) continue;
          float iw = b0 * t.v[0].w + b1 * t.v[1].w + b2 * t.v[2].w;
          float pw = 1.0f / iw;
          float at[9];
          for (int k = 0; k < 9; k++) at[k] = (b0 * t.a[0][k] + b1 * t.a[1][k] + b2 * t.a[2][k]) * pw;
          auto uv_at = [&](float x, float y, float &u, float &v) {
            float c0 = edge(t.v[1], t.v[2], x, y) * t.area_inv;
            float c1 = edge(t.v[2], t.v[0], x, y) * t.area_inv;
            float c2 = edge(t.v[0], t.v[1], x, y) * t.area_inv;
            float ww = 1.0f / (c0 * t.v[0].w + c1 * t.v[1].w + c2 * t.v[2].w);
            u = (c0 * t.a[0][0] + c1 * t.a[1][0] + c2 * t.a[2][0]) * ww;
            v = (c0 * t.a[0][1] + c1 * t.a[1][1] + c2 * t.a[2][1]) * ww;
          };
          V4 tex;
          if (view.is_mipmapped(at[1])) {
            float qx = (float)(px & ~1u) + 0.5f, qy = (float)(py & ~1u) + 0.5f;
            float u1, v1, u0, v0, u2, v2, u3, v3;
            uv_at(qx + 1.0f, fy, u1, v1);
            uv_at(qx, fy, u0, v0);
            uv_at(fx, qy + 1.0f, u2, v2);
            uv_at(fx, qy, u3, v3);
            tex = view.sample_block(at[0], at[1], u1 - u0, v1 - v0, u2 - u3, v2 - v3);
          } else {
            tex = view.nearest(at[0], at[1]);
          }
          lwo_gsky::V rgb{tex.x * at[2], tex.y * at[3], tex.z * at[4]};
          rgb = lwo_gsky::lerp(rgb, fog, fog_factor(sky, at[6], at[7], at[8]));
          if (t.translucent) {
            float a = tex.w * at[5];
            if (a <= 0.0f) continue;
            V4 d = unpack(color[li]);
            lwo_gsky::V mix = lwo_gsky::lerp(lwo_gsky::V{d.x, d.y, d.z}, rgb, a);
            color[li] = pack(mix.x, mix.y, mix.z);
          } else {
            if (tex.w < 0.5f) continue;
            depth[li] = z;
            color[li] = pack(rgb.x, rgb.y, rgb.z);
          }
        }
    }
    for (uint32_t k = 0; k < tw * th; k++) out[(size_t)(y0 + k / tw) * f.width + x0 + k % tw] = color[k];
  }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < threads; t++) pool.emplace_back(worker);
  worker();
  for (auto &t : pool) t.join();
  return out;
}

}  // namespace lwo_quick
