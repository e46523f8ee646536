// The host half of the game's sky (crates/lumenwright/src/render/sky.rs): the client's
// stars and their direction bins, QuickSky::new, celestial directions and the lightmap's
// colour, in f32 with libm's sinf, cosf, asinf and atan2f, the operations glam 0.34.1
// performs for the reference's Vec3 and Mat3 (column-major matrices; Mat3 * Vec3 is
// x_axis * v.x + y_axis * v.y + z_axis * v.z; normalize multiplies by 1 / length; cross
// is (y b.z - b.y z, z b.x - b.z x, x b.y - b.x y)). The device draws the maps from these
// (lw_hip_game_sky_maps).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "lw/host.h"

namespace {

constexpr float PI = 3.14159265358979323846f;
constexpr float TAU = 6.28318530717958647692f;
constexpr float FRAC_PI_2 = 1.57079632679489661923f;
constexpr float RADS_PER_DEG = PI / 180.0f;  // f32::to_radians

struct Vec3 {
  float x, y, z;
};
inline Vec3 vec(float x, float y, float z) { return Vec3{x, y, z}; }
inline Vec3 operator+(Vec3 a, Vec3 b) { return vec(a.x + b.x, a.y + b.y, a.z + b.z); }
inline Vec3 operator-(Vec3 a, Vec3 b) { return vec(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec3 operator*(Vec3 a, float s) { return vec(a.x * s, a.y * s, a.z * s); }
inline Vec3 operator/(Vec3 a, float s) { return vec(a.x / s, a.y / s, a.z / s); }
inline Vec3 operator-(Vec3 a) { return vec(-a.x, -a.y, -a.z); }
inline float dot(Vec3 a, Vec3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
inline float length(Vec3 a) { return std::sqrt(dot(a, a)); }
inline Vec3 normalize(Vec3 a) { return a * (1.0f / length(a)); }
inline Vec3 cross(Vec3 a, Vec3 b) {
  return vec(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
inline float fmax_r(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a > b ? a : b)); }
inline float fmin_r(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a < b ? a : b)); }

struct Mat3 {
  Vec3 x, y, z;  // columns
  Vec3 operator*(Vec3 v) const { return x * v.x + y * v.y + z * v.z; }
  Mat3 operator*(const Mat3 &m) const { return Mat3{*this * m.x, *this * m.y, *this * m.z}; }
  Mat3 transpose() const {
    return Mat3{vec(x.x, y.x, z.x), vec(x.y, y.y, z.y), vec(x.z, y.z, z.z)};
  }
  Vec3 row(int i) const {
    const float *c[3] = {&x.x, &y.x, &z.x};
    return vec(c[0][i], c[1][i], c[2][i]);
  }
};
Mat3 rotation_x(float a) {
  const float s = std::sin(a), c = std::cos(a);
  return Mat3{vec(1, 0, 0), vec(0, c, s), vec(0, -s, c)};
}
Mat3 rotation_y(float a) {
  const float s = std::sin(a), c = std::cos(a);
  return Mat3{vec(c, 0, -s), vec(0, 1, 0), vec(s, 0, c)};
}
Mat3 rotation_z(float a) {
  const float s = std::sin(a), c = std::cos(a);
  return Mat3{vec(c, s, 0), vec(-s, c, 0), vec(0, 0, 1)};
}

// sky.rs:54-57
Mat3 celestial_rotation(float angle_deg) { return rotation_y(-FRAC_PI_2) * rotation_x(angle_deg * RADS_PER_DEG); }

// java.util.Random's 48-bit LCG (sky.rs:72-98).
struct LegacyRandom {
  uint64_t s;
  static constexpr uint64_t MASK = (1ull << 48) - 1;
  explicit LegacyRandom(int64_t seed) : s(((uint64_t)seed ^ 0x5DEECE66Dull) & MASK) {}
  int32_t next(uint32_t bits) {
    s = (s * 0x5DEECE66Dull + 0xBull) & MASK;
    return (int32_t)(int64_t)(s >> (48 - bits));
  }
  float next_float() { return (float)next(24) * 5.9604645e-8f; }
  double next_double() {
    const int64_t hi = next(26), lo = next(27);
    return (double)((float)((hi << 27) + lo) * 1.110223e-16f);
  }
};

struct Star {
  Vec3 center, a1, a2;
};

// stars (sky.rs:109-137): the 1500 draws of SkyRenderer.buildStars, seed 10842.
std::vector<Star> make_stars() {
  LegacyRandom r(10842);
  std::vector<Star> out;
  for (int n = 0; n < 1500; n++) {
    const float x = r.next_float() * 2.0f - 1.0f;
    const float y = r.next_float() * 2.0f - 1.0f;
    const float z = r.next_float() * 2.0f - 1.0f;
    const float size = 0.15f + r.next_float() * 0.1f;
    const float len2 = x * x + y * y + z * z;
    if (len2 <= 0.010000001f || len2 >= 1.0f) continue;
    const Vec3 v = vec(x, y, z);
    const Vec3 center = v * (100.0f / length(v));
    const float theta = (float)(r.next_double() * (double)PI * 2.0);
    const Vec3 ndir = normalize(-center);
    const Vec3 left = normalize(cross(vec(0, 1, 0), ndir));
    const Vec3 upn = cross(ndir, left);
    const float s = std::sin(-theta), c = std::cos(-theta);
    out.push_back(Star{center, (left * c + upn * s) * size, (left * -s + upn * c) * size});
  }
  return out;
}

// body_quad (sky.rs:300-316)
void body_quad(Vec3 dir, float size, const float rect[4], bool flip, float out[4][4]) {
  const Vec3 c = dir * 100.0f;
  Vec3 ex = vec(0, 0, 1) - dir * dir.z;
  if (dot(ex, ex) < 1e-8f) ex = vec(1, 0, 0);
  ex = normalize(ex);
  const Vec3 ez = cross(ex, dir);
  const float s = flip ? -1.0f / size : 1.0f / size;
  const Vec3 u = ex * s, v = ez * s;
  const float rows[4][4] = {{c.x, c.y, c.z, 1.0f}, {u.x, u.y, u.z, 0.0f}, {v.x, v.y, v.z, 0.0f},
                            {rect[0], rect[1], rect[2], rect[3]}};
  std::memcpy(out, rows, sizeof(rows));
}

void put(float dst[4], Vec3 v, float w) {
  dst[0] = v.x;
  dst[1] = v.y;
  dst[2] = v.z;
  dst[3] = w;
}

}  // namespace

extern "C" {

void lw_celestial_direction(float angle_deg, float out[3]) {
  const Vec3 d = celestial_rotation(angle_deg) * vec(0, 1, 0);
  out[0] = d.x;
  out[1] = d.y;
  out[2] = d.z;
}

void lw_quick_sky_new(const lw_camera_rays *cam, uint32_t width, uint32_t height,
                      const lw_sky_description *d, lw_quick_sky *q) {
  std::memset(q, 0, sizeof(*q));
  put(q->sky_color, vec(d->sky[0], d->sky[1], d->sky[2]), d->sky_fog_end);
  put(q->fog_color, vec(d->fog[0], d->fog[1], d->fog[2]), d->render_fog[0]);
  q->fog_dist[0] = d->render_fog[1];
  q->fog_dist[1] = d->env_fog[0];
  q->fog_dist[2] = d->env_fog[1];
  q->fog_dist[3] = d->star_brightness * d->star_brightness;
  std::memcpy(q->sunrise, d->sunrise, sizeof(q->sunrise));
  q->flags[0] = d->sky_disc ? 1.0f : 0.0f;
  q->flags[1] = (d->sky_disc && d->sunrise[3] > 0.001f) ? 1.0f : 0.0f;
  q->flags[2] = (d->sky_disc && d->star_brightness > 0.0f) ? 1.0f : 0.0f;
  put(q->cam_forward, vec(cam->forward[0], cam->forward[1], cam->forward[2]), 0.0f);
  put(q->cam_right, vec(cam->right[0], cam->right[1], cam->right[2]), 0.0f);
  put(q->cam_up, vec(cam->up[0], cam->up[1], cam->up[2]), 0.0f);
  q->cam_misc[0] = cam->tan_half_fov;
  q->cam_misc[1] = cam->aspect;
  q->cam_misc[2] = (float)width;
  q->cam_misc[3] = (float)height;
  if (d->sky_disc) {
    if (d->has_sun) body_quad(vec(d->sun_dir[0], d->sun_dir[1], d->sun_dir[2]), 30.0f, d->sun_rect, false, q->sun);
    if (d->has_moon) body_quad(vec(d->moon_dir[0], d->moon_dir[1], d->moon_dir[2]), 20.0f, d->moon_rect, true, q->moon);
  }
  const Mat3 inv = celestial_rotation(d->star_angle).transpose();
  for (int i = 0; i < 3; i++) put(q->star_rot[i], inv.row(i), 0.0f);
  // renderSunriseAndSunset: XP(90 degrees), ZP(flip + 90 degrees), the fan scaled by the
  // glow's alpha along z (sky.rs:354-370).
  const float flip = std::sin(d->sun_angle * RADS_PER_DEG) < 0.0f ? 180.0f : 0.0f;
  const Mat3 diag{vec(1, 0, 0), vec(0, 1, 0), vec(0, 0, d->sunrise[3])};
  const Mat3 m = rotation_x(FRAC_PI_2) * rotation_z((flip + 90.0f) * RADS_PER_DEG) * diag;
  put(q->fan[0], m * vec(0.0f, 100.0f, 0.0f), d->sunrise[3]);
  for (int k = 0; k <= 16; k++) {
    const float t = (float)k * TAU / 16.0f;
    const float s = std::sin(t), c = std::cos(t);
    put(q->fan[k + 1], m * vec(s * 120.0f, c * 120.0f, -c * 40.0f), 0.0f);
  }
}

long lw_star_table(uint32_t *out, size_t cap) {
  static const std::vector<uint32_t> table = [] {
    const std::vector<Star> stars = make_stars();
    const int AZ = (int)LW_STAR_BINS_AZ, EL = (int)LW_STAR_BINS_EL;
    std::vector<std::vector<size_t>> bins((size_t)AZ * EL);
    const float da = TAU / (float)AZ, de = PI / (float)EL;
    for (size_t i = 0; i < stars.size(); i++) {
      const Star &s = stars[i];
      const Vec3 dir = s.center / 100.0f;
      const float radius = 2.0f * (length(s.a1) + length(s.a2)) / 100.0f;
      const float el = std::asin(dir.y);
      const float az = std::atan2(dir.z, dir.x);
      const float el0 = el - radius, el1 = el + radius;
      const int ie0 = std::max((int)std::floor((el0 + FRAC_PI_2) / de), 0);
      const int ie1 = std::min((int)std::floor((el1 + FRAC_PI_2) / de), EL - 1);
      const float cos_min = std::cos(fmin_r(fmax_r(std::fabs(el0), std::fabs(el1)), FRAC_PI_2));
      const float half = cos_min < 0.05f ? PI : radius / cos_min;
      const int ia0 = (int)std::floor((az - half + PI) / da);
      const int ia1 = (int)std::floor((az + half + PI) / da);
      for (int ie = ie0; ie <= ie1; ie++) {
        if (half >= PI) {
          for (int ia = 0; ia < AZ; ia++) bins[(size_t)ie * AZ + ia].push_back(i);
          continue;
        }
        for (int ia = ia0; ia <= ia1; ia++) {
          const int w = ((ia % AZ) + AZ) % AZ;
          std::vector<size_t> &b = bins[(size_t)ie * AZ + w];
          if (std::find(b.begin(), b.end(), i) == b.end()) b.push_back(i);
        }
      }
    }
    std::vector<uint32_t> t;
    uint32_t n = 0;
    for (auto &b : bins) {
      t.push_back(n);
      n += (uint32_t)b.size();
    }
    t.push_back(n);
    auto bits = [](float f) {
      uint32_t u;
      std::memcpy(&u, &f, 4);
      return u;
    };
    for (auto &b : bins)
      for (size_t i : b) {
        const Star &s = stars[i];
        const Vec3 a1 = s.a1 / dot(s.a1, s.a1), a2 = s.a2 / dot(s.a2, s.a2);
        for (const Vec3 &v : {s.center, a1, a2}) {
          t.push_back(bits(v.x));
          t.push_back(bits(v.y));
          t.push_back(bits(v.z));
        }
      }
    return t;
  }();
  if (out && cap >= table.size()) std::memcpy(out, table.data(), table.size() * 4);
  return (long)table.size();
}

void lw_disc_sector_trig(float s0[8], float c0[8], float s1[8], float c1[8]) {
  const float step = 0.785398163397448309616f;  // FRAC_PI_4
  for (int k = 0; k < 8; k++) {
    const float th0 = -PI + (float)k * step;
    s0[k] = std::sin(th0);
    c0[k] = std::cos(th0);
    s1[k] = std::sin(th0 + step);
    c1[k] = std::cos(th0 + step);
  }
}

void lw_lightmap_sky(float factor, const float colour[3], float brightness, float out[3]) {
  const float k = 0.75f * 0.04f;
  float c[3];
  for (int i = 0; i < 3; i++) {
    const float v = colour[i] * factor * 0.96f + k;  // mix75 (sky.rs:799)
    c[i] = fmin_r(fmax_r(v, 0.0f), 1.0f);            // glam clamp: max then min
  }
  const float mx = fmax_r(c[0], fmax_r(c[1], c[2]));
  float ng[3] = {c[0], c[1], c[2]};
  if (mx > 0.0f) {
    const float inv = 1.0f - mx;
    const float f = (1.0f - inv * inv * inv * inv) / mx;
    for (int i = 0; i < 3; i++) ng[i] = c[i] * f;
  }
  for (int i = 0; i < 3; i++) {
    const float v = (c[i] + (ng[i] - c[i]) * brightness) * 0.96f + k;
    out[i] = std::round(v * 255.0f) / 255.0f;
  }
}

}  // extern "C"
