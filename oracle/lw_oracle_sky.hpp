// ORACLE -- test infrastructure only. A restatement of the game's sky,
// crates/lumenwright/src/render/sky.rs: java.util.Random's LCG and the client's stars
// (:72-137), the star bins (:139-222), celestial rotations (:54-66), QuickSky::new
// (:320-373), QuickSky::background with its disc, glow, bodies and stars (:397-532),
// map_direction (:856-860), draw_map (:864-881), lightmap_sky (:798-815) and game_sky_maps
// (:906-953). f32 throughout with libm's sinf, cosf, asinf, atan2f and powf, as the reference
// computes on the CPU; Vec3 and Mat3 follow glam 0.34.1's published arithmetic (lerp is
// a * (1 - s) + b * s; clamp is max then min; Mat3 is column-major, M v = c0 v.x + c1 v.y +
// c2 v.z).
#pragma once

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <cstdint>
#include <cstring>
#include <vector>

namespace lwo_gsky {

constexpr float PI = 3.14159265358979323846f, TAU = 6.28318530717958647692f;
constexpr float FRAC_PI_2 = 1.57079632679489661923f, FRAC_PI_4 = 0.785398163397448309616f;
constexpr uint32_t BINS_AZ = 128, BINS_EL = 64, WORDS = 9;
constexpr uint32_t LENS_W = 2048, LENS_H = 1024, LIGHT_W = 360, LIGHT_H = 180;

struct V {
  float x, y, z;
  V operator+(V b) const { return {x + b.x, y + b.y, z + b.z}; }
  V operator-(V b) const { return {x - b.x, y - b.y, z - b.z}; }
  V operator*(float s) const { return {x * s, y * s, z * s}; }
  V operator/(float s) const { return {x / s, y / s, z / s}; }
  V operator-() const { return {-x, -y, -z}; }
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
static inline float dot(V a, V b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
static inline float len(V a) { return std::sqrt(dot(a, a)); }
static inline V unit(V a) { return a * (1.0f / len(a)); }
static inline V cross(V a, V b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
static inline float rmax(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a > b ? a : b)); }
static inline float rmin(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a < b ? a : b)); }
static inline float rclamp(float x, float lo, float hi) {  // f32::clamp
  if (x < lo) return lo;
  if (x > hi) return hi;
  return x;
}
static inline V vmin(V a, float b) { return {rmin(a.x, b), rmin(a.y, b), rmin(a.z, b)}; }
static inline V lerp(V a, V b, float s) { return a * (1.0f - s) + b * s; }
static inline V v4(const float *p) { return {p[0], p[1], p[2]}; }

struct M {  // columns
  V c0, c1, c2;
  V operator*(V v) const { return c0 * v.x + c1 * v.y + c2 * v.z; }
  M operator*(const M &m) const { return {*this * m.c0, *this * m.c1, *this * m.c2}; }
};
static inline M rot_x(float a) { float s = std::sin(a), c = std::cos(a); return {{1, 0, 0}, {0, c, s}, {0, -s, c}}; }
static inline M rot_y(float a) { float s = std::sin(a), c = std::cos(a); return {{c, 0, -s}, {0, 1, 0}, {s, 0, c}}; }
static inline M rot_z(float a) { float s = std::sin(a), c = std::cos(a); return {{c, s, 0}, {-s, c, 0}, {0, 0, 1}}; }
static inline float radians(float deg) { return deg * (PI / 180.0f); }

// sky.rs:54-66
static inline M celestial_rotation(float deg) { return rot_y(-FRAC_PI_2) * rot_x(radians(deg)); }
static inline V celestial_direction(float deg) { return celestial_rotation(deg) * V{0, 1, 0}; }

// sky.rs:72-98
struct Random {
  uint64_t s;
  explicit Random(int64_t seed) : s(((uint64_t)seed ^ 0x5DEECE66Dull) & ((1ull << 48) - 1)) {}
  int32_t next(int bits) {
    s = (s * 0x5DEECE66Dull + 0xBull) & ((1ull << 48) - 1);
    return (int32_t)(int64_t)(s >> (48 - bits));
  }
  float next_float() { return (float)next(24) * 5.9604645e-8f; }
  double next_double() {
    int64_t hi = next(26), lo = next(27);
    return (double)((float)((hi << 27) + lo) * 1.110223e-16f);
  }
};

struct Star {
  V center, a1, a2;
};

// sky.rs:109-137
static std::vector<Star> stars() {
  Random r(10842);
  std::vector<Star> out;
  for (int n = 0; n < 1500; n++) {
    float x = r.next_float() * 2.0f - 1.0f, y = r.next_float() * 2.0f - 1.0f, z = r.next_float() * 2.0f - 1.0f;
    float size = 0.15f + r.next_float() * 0.1f;
    float len2 = x * x + y * y + z * z;
    if (len2 <= 0.010000001f || len2 >= 1.0f) continue;
    V v{x, y, z};
    V center = v * (100.0f / len(v));
    float theta = (float)(r.next_double() * (double)PI * 2.0);
    V nd = unit(-center);
    V left = unit(cross(V{0, 1, 0}, nd));
    V upn = cross(nd, left);
    float s = std::sin(-theta), c = std::cos(-theta);
    out.push_back({center, (left * c + upn * s) * size, (left * -s + upn * c) * size});
  }
  return out;
}

// sky.rs:152-222
static std::vector<uint32_t> star_table() {
  const std::vector<Star> st = stars();
  const int NA = (int)BINS_AZ, NE = (int)BINS_EL;
  std::vector<std::vector<size_t>> bins((size_t)NA * NE);
  const float da = TAU / (float)NA, de = PI / (float)NE;
  for (size_t i = 0; i < st.size(); i++) {
    V dir = st[i].center / 100.0f;
    float radius = 2.0f * (len(st[i].a1) + len(st[i].a2)) / 100.0f;
    float el = std::asin(dir.y), az = std::atan2(dir.z, dir.x);
    float el0 = el - radius, el1 = el + radius;
    int ie0 = std::max((int)std::floor((el0 + FRAC_PI_2) / de), 0);
    int ie1 = std::min((int)std::floor((el1 + FRAC_PI_2) / de), NE - 1);
    float cos_min = std::cos(rmin(rmax(std::fabs(el0), std::fabs(el1)), FRAC_PI_2));
    float half = cos_min < 0.05f ? PI : radius / cos_min;
    int ia0 = (int)std::floor((az - half + PI) / da), ia1 = (int)std::floor((az + half + PI) / da);
    for (int ie = ie0; ie <= ie1; ie++) {
      if (half >= PI) {
        for (int ia = 0; ia < NA; ia++) bins[(size_t)ie * NA + ia].push_back(i);
        continue;
      }
      for (int ia = ia0; ia <= ia1; ia++) {
        auto &b = bins[(size_t)ie * NA + (size_t)(((ia % NA) + NA) % NA)];
        if (std::find(b.begin(), b.end(), i) == b.end()) b.push_back(i);
      }
    }
  }
  std::vector<uint32_t> out;
  uint32_t n = 0;
  for (auto &b : bins) {
    out.push_back(n);
    n += (uint32_t)b.size();
  }
  out.push_back(n);
  for (auto &b : bins)
    for (size_t i : b) {
      V a1 = st[i].a1 / dot(st[i].a1, st[i].a1), a2 = st[i].a2 / dot(st[i].a2, st[i].a2);
      for (V v : {st[i].center, a1, a2})
        for (int k = 0; k < 3; k++) {
          float f = v[k];
          uint32_t u;
          std::memcpy(&u, &f, 4);
          out.push_back(u);
        }
    }
  return out;
}

// SkyDescription (sky.rs:225-253) and QuickSky (sky.rs:256-283), the library's layouts.
struct Description {
  float sky[3], fog[3], sky_fog_end, render_fog[2], env_fog[2];
  uint32_t sky_disc;
  float sunrise[4], sun_angle;
  uint32_t has_sun;
  float sun_dir[3], sun_rect[4];
  uint32_t has_moon;
  float moon_dir[3], moon_rect[4], star_angle, star_brightness;
};
struct Quick {
  float sky_color[4], fog_color[4], fog_dist[4], sunrise[4], flags[4];
  float sun[4][4], moon[4][4], star_rot[3][4], fan[18][4];
  float cam_forward[4], cam_right[4], cam_up[4], cam_misc[4];
};
static_assert(sizeof(Quick) == 152 * 4, "QuickSky layout");

static void set4(float *d, V v, float w) {
  d[0] = v.x;
  d[1] = v.y;
  d[2] = v.z;
  d[3] = w;
}

// sky.rs:300-316
static void body_quad(V dir, float size, const float rect[4], bool flip, float out[4][4]) {
  V c = dir * 100.0f;
  V ex = V{0, 0, 1} - dir * dir.z;
  if (dot(ex, ex) < 1e-8f) ex = V{1, 0, 0};
  ex = unit(ex);
  V ez = cross(ex, dir);
  float s = flip ? -1.0f / size : 1.0f / size;
  set4(out[0], c, 1.0f);
  set4(out[1], ex * s, 0.0f);
  set4(out[2], ez * s, 0.0f);
  for (int k = 0; k < 4; k++) out[3][k] = rect[k];
}

// sky.rs:320-373 (the camera: forward, right, up, tan(fov/2), aspect)
static Quick quick_new(const float cam[11], uint32_t w, uint32_t h, const Description &d) {
  Quick q;
  std::memset(&q, 0, sizeof(q));
  set4(q.sky_color, v4(d.sky), d.sky_fog_end);
  set4(q.fog_color, v4(d.fog), d.render_fog[0]);
  q.fog_dist[0] = d.render_fog[1];
  q.fog_dist[1] = d.env_fog[0];
  q.fog_dist[2] = d.env_fog[1];
  q.fog_dist[3] = d.star_brightness * d.star_brightness;
  for (int k = 0; k < 4; k++) q.sunrise[k] = d.sunrise[k];
  q.flags[0] = d.sky_disc ? 1.0f : 0.0f;
  q.flags[1] = (d.sky_disc && d.sunrise[3] > 0.001f) ? 1.0f : 0.0f;
  q.flags[2] = (d.sky_disc && d.star_brightness > 0.0f) ? 1.0f : 0.0f;
  set4(q.cam_forward, v4(cam), 0.0f);
  set4(q.cam_right, v4(cam + 3), 0.0f);
  set4(q.cam_up, v4(cam + 6), 0.0f);
  q.cam_misc[0] = cam[9];
  q.cam_misc[1] = cam[10];
  q.cam_misc[2] = (float)w;
  q.cam_misc[3] = (float)h;
  if (d.sky_disc) {
    if (d.has_sun) body_quad(v4(d.sun_dir), 30.0f, d.sun_rect, false, q.sun);
    if (d.has_moon) body_quad(v4(d.moon_dir), 20.0f, d.moon_rect, true, q.moon);
  }
  M r = celestial_rotation(d.star_angle);  // its transpose's rows are its columns
  set4(q.star_rot[0], r.c0, 0.0f);
  set4(q.star_rot[1], r.c1, 0.0f);
  set4(q.star_rot[2], r.c2, 0.0f);
  float flip = std::sin(radians(d.sun_angle)) < 0.0f ? 180.0f : 0.0f;
  M m = rot_x(FRAC_PI_2) * rot_z(radians(flip + 90.0f)) * M{{1, 0, 0}, {0, 1, 0}, {0, 0, d.sunrise[3]}};
  set4(q.fan[0], m * V{0.0f, 100.0f, 0.0f}, d.sunrise[3]);
  for (int k = 0; k <= 16; k++) {
    float t = (float)k * TAU / 16.0f;
    float s = std::sin(t), c = std::cos(t);
    set4(q.fan[k + 1], m * V{s * 120.0f, c * 120.0f, -c * 40.0f}, 0.0f);
  }
  return q;
}

static float linear_fog(float v, float start, float end) {
  if (v <= start) return 0.0f;
  if (v >= end) return 1.0f;
  return (v - start) / (end - start);
}

// sky.rs:402-422
static bool disc(const Quick &q, V d, V &out) {
  if (d.y <= 0.0f) return false;
  V p = d * (16.0f / d.y);
  float phi = std::atan2(p.z, p.x);
  float k = rclamp(std::floor((phi + PI) / FRAC_PI_4), 0.0f, 7.0f);
  float th0 = -PI + k * FRAC_PI_4;
  float s0 = std::sin(th0), c0 = std::cos(th0), s1 = std::sin(th0 + FRAC_PI_4), c1 = std::cos(th0 + FRAC_PI_4);
  float r0x = 512.0f * c0, r0z = 512.0f * s0, r1x = 512.0f * c1, r1z = 512.0f * s1;
  float det = r0x * r1z - r1x * r0z;
  float b1 = (p.x * r1z - r1x * p.z) / det, b2 = (r0x * p.z - p.x * r0z) / det;
  float s = b1 + b2;
  if (s > 1.0f) return false;
  float end = q.sky_color[3];
  float sph = 16.0f * (1.0f - s) + 512.24994f * s;
  float cyl = 16.0f * (1.0f - s) + 512.0f * s;
  float f = rmax(linear_fog(sph, 0.0f, end), linear_fog(cyl, end, end));
  out = lerp(v4(q.sky_color), v4(q.fog_color), f);
  return true;
}

// sky.rs:425-454
static float fan_alpha(const Quick &q, V d) {
  V c = v4(q.fan[0]);
  float best = INFINITY, alpha = 0.0f;
  for (int k = 0; k < 16; k++) {
    V e1 = v4(q.fan[k + 1]) - c, e2 = v4(q.fan[k + 2]) - c;
    V pv = cross(d, e2);
    float det = dot(e1, pv);
    if (std::fabs(det) < 1e-6f) continue;
    float inv = 1.0f / det;
    V tv = -c;
    float u = dot(tv, pv) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) continue;
    V qv = cross(tv, e1);
    float v = dot(d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) continue;
    float t = dot(e2, qv) * inv;
    if (t > 0.0f && t < best) {
      best = t;
      alpha = (1.0f - u - v) * q.fan[0][3];
    }
  }
  return alpha;
}

// sky.rs:457-483
static void body(const float b[4][4], V d, const uint32_t *atlas, uint32_t atlas_w, float out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0.0f;
  if (b[0][3] == 0.0f || b[3][2] <= 0.0f) return;
  V c = v4(b[0]);
  float dn = dot(d, c);
  if (dn <= 0.0f) return;
  V q = d * (dot(c, c) / dn) - c;
  float lx = dot(q, v4(b[1])), lz = dot(q, v4(b[2]));
  if (std::fabs(lx) >= 1.0f || std::fabs(lz) >= 1.0f) return;
  float u = b[3][0] + (lx * 0.5f + 0.5f) * b[3][2];
  float v = b[3][1] + (lz * 0.5f + 0.5f) * b[3][3];
  auto as_u32 = [](float f) -> uint32_t { return f > 0.0f ? (f >= 4294967296.0f ? 0xFFFFFFFFu : (uint32_t)f) : 0u; };
  uint32_t t = atlas[as_u32(std::floor(v)) * atlas_w + as_u32(std::floor(u))];
  out[0] = (float)(t & 0xFF) / 255.0f;
  out[1] = (float)((t >> 8) & 0xFF) / 255.0f;
  out[2] = (float)((t >> 16) & 0xFF) / 255.0f;
  out[3] = (float)(t >> 24) / 255.0f;
}

// sky.rs:139-150, 486-507
static uint32_t star_hits(const Quick &q, V d, const std::vector<uint32_t> &table) {
  V qd{dot(v4(q.star_rot[0]), d), dot(v4(q.star_rot[1]), d), dot(v4(q.star_rot[2]), d)};
  float az = std::atan2(qd.z, qd.x);
  float el = std::asin(rclamp(qd.y, -1.0f, 1.0f));
  int ia = (int)std::floor(((az + PI) / TAU) * (float)BINS_AZ);
  int ie = (int)std::floor(((el + FRAC_PI_2) / PI) * (float)BINS_EL);
  ia = std::min(std::max(ia, 0), (int)BINS_AZ - 1);
  ie = std::min(std::max(ie, 0), (int)BINS_EL - 1);
  size_t b = (size_t)ie * BINS_AZ + (size_t)ia, base = (size_t)BINS_AZ * BINS_EL + 1;
  auto f = [&](size_t w) {
    float x;
    std::memcpy(&x, &table[w], 4);
    return x;
  };
  uint32_t n = 0;
  for (uint32_t e = table[b]; e < table[b + 1]; e++) {
    size_t o = base + (size_t)e * WORDS;
    V c{f(o), f(o + 1), f(o + 2)};
    float dn = dot(qd, c);
    if (dn <= 0.0f) continue;
    V p = qd * (dot(c, c) / dn) - c;
    if (std::fabs(dot(p, V{f(o + 3), f(o + 4), f(o + 5)})) < 1.0f &&
        std::fabs(dot(p, V{f(o + 6), f(o + 7), f(o + 8)})) < 1.0f)
      n++;
  }
  return n;
}

// sky.rs:509-532
static V background(const Quick &q, V d, const uint32_t *atlas, uint32_t atlas_w, const std::vector<uint32_t> &table) {
  V c = v4(q.fog_color);
  V dc;
  if (q.flags[0] > 0.5f && disc(q, d, dc)) c = dc;
  if (q.flags[1] > 0.5f) c = lerp(c, v4(q.sunrise), fan_alpha(q, d));
  for (const auto *b : {&q.sun, &q.moon}) {
    float t[4];
    body(*b, d, atlas, atlas_w, t);
    c = vmin(c + V{t[0], t[1], t[2]} * t[3], 1.0f);
  }
  if (q.flags[2] > 0.5f) {
    uint32_t n = star_hits(q, d, table);
    for (uint32_t i = 0; i < n; i++) c = vmin(c + V{q.fog_dist[3], q.fog_dist[3], q.fog_dist[3]}, 1.0f);
  }
  return {rmin(rmax(c.x, 0.0f), 1.0f), rmin(rmax(c.y, 0.0f), 1.0f), rmin(rmax(c.z, 0.0f), 1.0f)};
}

// sky.rs:856-860
static V map_direction(uint32_t i, uint32_t j, uint32_t w, uint32_t h) {
  float az = ((float)i + 0.5f) / (float)w * TAU;
  float el = ((float)j + 0.5f) / (float)h * PI - FRAC_PI_2;
  return {std::cos(el) * std::cos(az), std::sin(el), std::cos(el) * std::sin(az)};
}

// sky.rs:864-881: c^2.2 per channel (math.rs GAMMA); the texels on `threads` threads, rows
// taken dynamically, as the reference maps them with rayon (:869-872).
static void draw_map(const Quick &q, uint32_t w, uint32_t h, const uint32_t *atlas, uint32_t atlas_w,
                     const std::vector<uint32_t> &table, float *out, unsigned threads = 1) {
  std::atomic<uint32_t> next_row{0};
  auto worker = [&]() {
    for (uint32_t j = next_row.fetch_add(1); j < h; j = next_row.fetch_add(1))
      for (uint32_t k = j * w; k < (j + 1) * w; k++) {
        V c = background(q, map_direction(k % w, k / w, w, h), atlas, atlas_w, table);
        out[3 * k] = std::pow(c.x, 2.2f);
        out[3 * k + 1] = std::pow(c.y, 2.2f);
        out[3 * k + 2] = std::pow(c.z, 2.2f);
      }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < threads; t++) pool.emplace_back(worker);
  worker();
  for (auto &t : pool) t.join();
}

// sky.rs:798-815
static V lightmap_sky(float factor, V colour, float brightness) {
  auto mix75 = [](V c) { return c * 0.96f + V{0.75f * 0.04f, 0.75f * 0.04f, 0.75f * 0.04f}; };
  V m = mix75(colour * factor);
  V c{rmin(rmax(m.x, 0.0f), 1.0f), rmin(rmax(m.y, 0.0f), 1.0f), rmin(rmax(m.z, 0.0f), 1.0f)};
  float mx = rmax(c.x, rmax(c.y, c.z));
  V ng = c;
  if (mx > 0.0f) {
    float inv = 1.0f - mx;
    ng = c * ((1.0f - inv * inv * inv * inv) / mx);
  }
  V r = mix75(c + (ng - c) * brightness);
  return {std::round(r.x * 255.0f) / 255.0f, std::round(r.y * 255.0f) / 255.0f, std::round(r.z * 255.0f) / 255.0f};
}

// sky.rs:906-953. out: the lens's map then the light's; info: light, mean, scale, veil.
static void game_sky_maps(const Quick &lens, const Quick &light, float factor, V colour, const uint32_t *atlas,
                          uint32_t atlas_w, const std::vector<uint32_t> &table, float *out, float info[10],
                          unsigned threads = 1) {
  float *seen = out, *game = out + (size_t)LENS_W * LENS_H * 3;
  draw_map(lens, LENS_W, LENS_H, atlas, atlas_w, table, seen, threads);
  draw_map(light, LIGHT_W, LIGHT_H, atlas, atlas_w, table, game, threads);
  double acc[3] = {0, 0, 0}, weight = 0.0;
  for (uint32_t k = 0; k < LIGHT_W * LIGHT_H; k++) {
    V d = map_direction(k % LIGHT_W, k / LIGHT_W, LIGHT_W, LIGHT_H);
    if (d.y > 0.0f) {
      double w = (double)(d.y * std::sqrt(rmax(1.0f - d.y * d.y, 0.0f)));
      for (int c = 0; c < 3; c++) acc[c] += (double)game[3 * k + c] * w;
      weight += w;
    }
  }
  V mean{(float)(acc[0] / weight), (float)(acc[1] / weight), (float)(acc[2] / weight)};
  V t = lightmap_sky(factor, colour, 0.5f);
  V target{std::pow(t.x, 2.2f), std::pow(t.y, 2.2f), std::pow(t.z, 2.2f)};
  float scale = 50.0f;
  for (int k = 0; k < 3; k++)
    if (mean[k] > 0.0f) scale = rmin(scale, target[k] / mean[k]);
  V vr = target / scale - mean;
  V veil{rmax(vr.x, 0.0f), rmax(vr.y, 0.0f), rmax(vr.z, 0.0f)};
  for (uint32_t k = 0; k < LIGHT_W * LIGHT_H; k++) {
    V g{game[3 * k], game[3 * k + 1], game[3 * k + 2]};
    V o = (g + veil) * scale;
    game[3 * k] = o.x;
    game[3 * k + 1] = o.y;
    game[3 * k + 2] = o.z;
  }
  const float vals[10] = {target.x, target.y, target.z, mean.x, mean.y, mean.z, scale, veil.x, veil.y, veil.z};
  std::memcpy(info, vals, sizeof(vals));
}

}  // namespace lwo_gsky
