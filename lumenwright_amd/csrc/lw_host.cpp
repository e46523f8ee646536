// Host-side frame preparation: the sun, the Preetham sky and its cache, the camera, the
// kernel's Params, and the synthetic worlds laid out as the reference's build_hq lays out a
// loaded world. See include/lw/host.h.

#include "lw_host.hpp"

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lw_math.hpp"
#include "lw_sky.hpp"

using namespace lw;

namespace lw {

void gamma_lut(float out[256]) {
  for (uint32_t k = 0; k < 256; k++) out[k] = gamma_to_linear((float)k / 255.0f);
}

}  // namespace lw

namespace {

constexpr float SUN_RADIUS = 0.03f;                // sky.rs:565
constexpr float SUN_TEXTURE_BRIGHTNESS = 16.338118f;  // sky.rs:571
constexpr float NIGHT_SKY_LIGHT = 0.03f;           // sky.rs:617
const float CIVIL_TWILIGHT_END = -6.0f * PI_F / 180.0f;  // sky.rs:613

// glam's f32 Vec3::normalize: self * (1 / length).
inline void normalize3(float v[3]) {
  float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  float r = 1.0f / l;
  v[0] *= r;
  v[1] *= r;
  v[2] *= r;
}

// Rust f64::rem_euclid.
inline double rem_euclid(double a, double b) {
  double r = std::fmod(a, b);
  return r < 0.0 ? r + std::fabs(b) : r;
}

// ---- synthetic worlds ----

inline uint64_t mix64(uint64_t x) {  // SplitMix64 finaliser
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
inline uint64_t hash5(uint64_t seed, uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
  return mix64(mix64(mix64(mix64(mix64(seed) ^ a) ^ b) ^ c) ^ d);
}
inline double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

enum : uint64_t {
  H_LATTICE = 1, H_TYPE, H_EMIT, H_PLANT, H_PLANT_TYPE, H_TEXEL, H_ALPHA, H_EMITTANCE, H_EMIT_COLUMN
};

double value_noise(uint64_t seed, int32_t x, int32_t z) {
  static const int lattice[4] = {32, 16, 8, 4};
  static const double gain[4] = {1.0, 0.5, 0.25, 0.125};
  double sum = 0.0, norm = 0.0;
  for (int o = 0; o < 4; o++) {
    const double L = lattice[o];
    const double fx = x / L, fz = z / L;
    const double ix = std::floor(fx), iz = std::floor(fz);
    const double tx = fx - ix, tz = fz - iz;
    const double sx = tx * tx * (3.0 - 2.0 * tx), sz = tz * tz * (3.0 - 2.0 * tz);
    auto r = [&](double i, double j) {
      return 2.0 * u01(hash5(seed, H_LATTICE, (uint64_t)o, (uint64_t)(int64_t)i, (uint64_t)(int64_t)j)) - 1.0;
    };
    const double a = r(ix, iz) + sx * (r(ix + 1, iz) - r(ix, iz));
    const double b = r(ix, iz + 1) + sx * (r(ix + 1, iz + 1) - r(ix, iz + 1));
    sum += gain[o] * (a + sz * (b - a));
    norm += gain[o];
  }
  return sum / norm;
}

int32_t terrain_height(uint64_t seed, int32_t x, int32_t z) {
  return (int32_t)std::floor(64.0 + 24.0 * value_noise(seed, x, z));
}

lw_quad make_quad(const float p0[3], const float eu[3], const float ev[3], uint32_t flags,
                  float emittance, uint32_t sprite_x, uint32_t sprite_y) {
  lw_quad q;
  std::memset(&q, 0, sizeof(q));
  for (int c = 0; c < 3; c++) {
    q.p0[c] = p0[c];
    q.eu[c] = eu[c];
    q.ev[c] = ev[c];
  }
  q.p0[3] = bits_f(flags);
  q.eu[3] = bits_f(0xFFFFFFFFu);
  q.ev[3] = emittance;
  q.uv[0] = (float)sprite_x;
  q.uv[1] = (float)sprite_y;
  q.uv[2] = 16.0f;
  q.uv[3] = 0.0f;
  q.uvt[0] = 0.0f;
  q.uvt[1] = 16.0f;
  return q;
}

// The unit cube's faces in block-local space, outward winding (normal = ev x eu), in the
// order of the reference's test cube (north, south, east, west, up, down).
void cube_faces(std::vector<lw_quad> &quads, uint32_t flags, uint32_t sx, uint32_t sy,
                const float emittance[6]) {
  static const float P[6][3] = {{1, 1, 0}, {0, 1, 1}, {1, 1, 1}, {0, 1, 0}, {0, 1, 0}, {0, 0, 1}};
  static const float U[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, 0, -1}, {0, 0, 1}, {1, 0, 0}, {1, 0, 0}};
  static const float V[6][3] = {{0, -1, 0}, {0, -1, 0}, {0, -1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  for (int f = 0; f < 6; f++) quads.push_back(make_quad(P[f], U[f], V[f], flags, emittance[f], sx, sy));
}

// A cross plant: two diagonal planes, each met from both sides by a pair of opposite
// front-only quads, alpha-tested.
void cross_faces(std::vector<lw_quad> &quads, uint32_t sx, uint32_t sy) {
  const uint32_t flags = LW_ALPHA_CUTOUT | LW_FLAG_FRONT_ONLY;
  static const float P[4][3] = {{0, 1, 0}, {1, 1, 1}, {0, 1, 1}, {1, 1, 0}};
  static const float U[4][3] = {{1, 0, 1}, {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}};
  static const float V[3] = {0, -1, 0};
  for (int f = 0; f < 4; f++) quads.push_back(make_quad(P[f], U[f], V, flags, 0.0f, sx, sy));
}

// voxel.rs:532 (EMITTER_CELL) and emitter_list (voxel.rs:542-607).
void emitter_list(const std::vector<lw_model> &models, const std::vector<std::pair<std::array<int32_t, 3>, uint32_t>> &sites,
                  std::vector<lw_emitter> &emitters, std::vector<uint32_t> &grid) {
  const int32_t CELL = 10;
  emitters.clear();
  grid.clear();
  if (sites.empty()) {
    grid.assign(LW_EMITTER_GRID_HEADER, 0u);
    grid[0] = (uint32_t)CELL;
    emitters.push_back(lw_emitter{});
    return;
  }
  for (auto &s : sites) {
    const lw_model &m = models[s.second];
    lw_emitter e;
    std::memset(&e, 0, sizeof(e));
    e.pos[0] = (float)s.first[0];
    e.pos[1] = (float)s.first[1];
    e.pos[2] = (float)s.first[2];
    e.pos[3] = bits_f(m.quad_offset);
    e.faces[0] = m.quad_count;
    emitters.push_back(e);
  }
  auto div_euclid = [](int32_t a, int32_t b) {
    int32_t q = a / b;
    if (a % b < 0) q -= 1;
    return q;
  };
  int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  for (auto &s : sites)
    for (int a = 0; a < 3; a++) {
      int32_t g = div_euclid(s.first[a], CELL);
      lo[a] = std::min(lo[a], g);
      hi[a] = std::max(hi[a], g);
    }
  int32_t offset[3], size[3];
  for (int a = 0; a < 3; a++) {
    offset[a] = lo[a] - 1;
    size[a] = hi[a] - lo[a] + 3;
  }
  auto index = [&](int32_t gx, int32_t gy, int32_t gz) -> size_t {
    int32_t l0 = gx - offset[0], l1 = gy - offset[1], l2 = gz - offset[2];
    return (size_t)((l1 * size[2] + l2) * size[0] + l0);
  };
  std::vector<std::vector<uint32_t>> cells((size_t)size[0] * size[1] * size[2]);
  for (size_t i = 0; i < sites.size(); i++) {
    int32_t g[3];
    for (int a = 0; a < 3; a++) g[a] = div_euclid(sites[i].first[a], CELL);
    for (int dy = -1; dy <= 1; dy++)
      for (int dz = -1; dz <= 1; dz++)
        for (int dx = -1; dx <= 1; dx++) cells[index(g[0] + dx, g[1] + dy, g[2] + dz)].push_back((uint32_t)i);
  }
  grid = {(uint32_t)CELL, (uint32_t)size[0], (uint32_t)size[1], (uint32_t)size[2],
          (uint32_t)offset[0], (uint32_t)offset[1], (uint32_t)offset[2], 0u};
  std::vector<uint32_t> list;
  for (auto &c : cells) {
    grid.push_back((uint32_t)list.size());
    grid.push_back((uint32_t)c.size());
    list.insert(list.end(), c.begin(), c.end());
  }
  grid.insert(grid.end(), list.begin(), list.end());
}

template <typename T>
T *to_heap(const std::vector<T> &v) {
  T *p = (T *)std::malloc(std::max<size_t>(1, v.size()) * sizeof(T));
  if (!p) throw std::bad_alloc();
  if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

thread_local std::string g_host_error;

}  // namespace

extern "C" {

float lw_celestial_angle(int64_t ticks) {
  int64_t t = ticks % 24000;
  if (t < 0) t += 24000;
  double f = rem_euclid((double)t / 24000.0 - 0.25, 1.0);
  double g = 0.5 - std::cos(f * M_PI) / 2.0;
  return (float)((f * 2.0 + g) / 3.0);
}

void lw_sun_direction_from_angle(float angle, float out[3]) {
  const float TAU_F = 6.28318530717958647692f;
  float t = angle * TAU_F;
  out[0] = -std::sin(t);
  out[1] = std::cos(t);
  out[2] = 0.0f;
  normalize3(out);
}

void lw_sun_direction(float altitude_deg, float azimuth_deg, float out[3]) {
  const float RADS = PI_F / 180.0f;  // f32::to_radians
  float alt = altitude_deg * RADS, az = azimuth_deg * RADS;
  out[0] = std::sin(az) * std::cos(alt);
  out[1] = std::sin(alt);
  out[2] = -std::cos(az) * std::cos(alt);
  normalize3(out);
}

float lw_twilight(float altitude) {
  float x = (altitude - CIVIL_TWILIGHT_END) / (-SUN_RADIUS - CIVIL_TWILIGHT_END);
  x = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
  return std::fmax(x, NIGHT_SKY_LIGHT);
}

void lw_hq_sky_new(const float sun_dir[3], float sun_intensity, float sky_intensity, lw_hq_sky *out) {
  float y = sun_dir[1] < -1.0f ? -1.0f : (sun_dir[1] > 1.0f ? 1.0f : sun_dir[1]);
  float altitude = std::asin(y);
  float c = std::pow(sun_intensity, GAMMA);
  for (int k = 0; k < 3; k++) {
    out->sun_dir[k] = sun_dir[k];
    out->sun_color[k] = c;
  }
  out->sky_light = sky_intensity * lw_twilight(altitude);
}

void lw_preetham_new(const float sun_dir[3], double TURB, lw_preetham *out) {
  static const double X_ZENITH[3][4] = {{0.00166, -0.00375, 0.00209, 0.0},
                                        {-0.02903, 0.06377, -0.03203, 0.00394},
                                        {0.11693, -0.21196, 0.06052, 0.25886}};
  static const double Y_ZENITH[3][4] = {{0.00275, -0.00610, 0.00317, 0.0},
                                        {-0.04214, 0.08970, -0.04153, 0.00516},
                                        {0.15346, -0.26756, 0.06670, 0.26688}};
  static const double MDX[5][2] = {{-0.0193, -0.2592}, {-0.0665, 0.0008}, {-0.0004, 0.2125},
                                   {-0.0641, -0.8989}, {-0.0033, 0.0452}};
  static const double MDY[5][2] = {{-0.0167, -0.2608}, {-0.0950, 0.0092}, {-0.0079, 0.2102},
                                   {-0.0441, -1.6537}, {-0.0109, 0.0529}};
  static const double MD_LUM[5][2] = {{0.1787, -1.4630}, {-0.3554, 0.4275}, {-0.0227, 5.3251},
                                      {0.1206, -2.5771}, {-0.0670, 0.3703}};
  // glam DVec3::normalize: self * (1 / length)
  double d[3] = {sun_dir[0], sun_dir[1], sun_dir[2]};
  double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double rl = 1.0 / len;
  for (double &c : d) c *= rl;
  double azimuth = std::atan2(d[2], d[0]);
  double yc = d[1] < -1.0 ? -1.0 : (d[1] > 1.0 ? 1.0 : d[1]);
  double alt = std::asin(yc);
  alt = alt < 0.0 ? 0.0 : (alt > M_PI ? M_PI : alt);
  if (alt > M_PI_2) alt = M_PI - alt;
  double r = std::fabs(std::cos(alt));
  double sun[3] = {std::cos(azimuth) * r, std::sin(alt), std::sin(azimuth) * r};

  double cx[5], cy[5], cl[5];
  for (int i = 0; i < 5; i++) {
    cx[i] = MDX[i][0] * TURB + MDX[i][1];
    cy[i] = MDY[i][0] * TURB + MDY[i][1];
    cl[i] = MD_LUM[i][0] * TURB + MD_LUM[i][1];
  }
  auto perez_f = [](double cos_theta, double gamma, double cos2_gamma, const double *c) {
    return (1.0 + c[0] * std::exp(c[1] / cos_theta)) *
           (1.0 + c[2] * std::exp(c[3] * gamma) + c[4] * cos2_gamma);
  };
  auto chroma = [&](double st, const double m[3][4]) {
    double t1 = st, t2 = st * st, t3 = st * st * st;
    auto row = [&](const double *rr) { return rr[0] * t3 + rr[1] * t2 + rr[2] * t1 + rr[3]; };
    return TURB * TURB * row(m[0]) + TURB * row(m[1]) + row(m[2]);
  };
  double sun_theta = M_PI_2 - alt;
  double cst = std::cos(sun_theta);
  double cos2_theta = cst * cst;  // powi(2)
  double chi = (4.0 / 9.0 - TURB / 120.0) * (M_PI - 2.0 * sun_theta);
  double zenith_lum = std::fabs((4.0453 * TURB - 4.9710) * std::tan(chi) - 0.2155 * TURB + 2.4192);
  double zenith_x = chroma(sun_theta, X_ZENITH);
  double zenith_y = chroma(sun_theta, Y_ZENITH);
  auto f0 = [&](const double *c) { return 1.0 / perez_f(1.0, sun_theta, cos2_theta, c); };
  for (int k = 0; k < 3; k++) out->sun[k] = (float)sun[k];
  out->zenith[0] = (float)(zenith_x * f0(cx));
  out->zenith[1] = (float)(zenith_y * f0(cy));
  out->zenith[2] = (float)(zenith_lum * f0(cl));
  for (int i = 0; i < 5; i++) {
    out->perez[i][0] = (float)cx[i];
    out->perez[i][1] = (float)cy[i];
    out->perez[i][2] = (float)cl[i];
  }
}

void lw_sun_texture_axes(const float sw_in[3], float su_out[3], float sv_out[3]) {
  double sw[3] = {sw_in[0], sw_in[1], sw_in[2]};
  double su[3] = {0.0, 0.0, 0.0};
  if (std::fabs(sw[0]) > 0.1)
    su[1] = 1.0;
  else
    su[0] = 1.0;
  auto cross = [](const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - b[1] * a[2];
    o[1] = a[2] * b[0] - b[2] * a[0];
    o[2] = a[0] * b[1] - b[0] * a[1];
  };
  double sv[3];
  cross(sw, su, sv);
  double l = std::sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
  double rl = 1.0 / l;
  for (double &c : sv) c *= rl;
  cross(sv, sw, su);
  for (int k = 0; k < 3; k++) {
    su_out[k] = (float)su[k];
    sv_out[k] = (float)sv[k];
  }
}

long lw_sky_cache(const lw_params *params, float *out, size_t cap) {
  const uint32_t n = as_u32(params->misc_f[2]);
  if (n == 0) return 0;
  const uint32_t m = n + 1;
  const size_t need = (size_t)m * m * 3;
  if (cap < need) return -1;
  for (uint32_t i = 0; i < m; i++)
    for (uint32_t j = 0; j < m; j++) lw::sky_cache_entry(*params, n, i, j, out + (size_t)(i * m + j) * 3);
  return (long)need;
}

void lw_gamma_lut(float out[256]) { lw::gamma_lut(out); }

void lw_map_trig(uint32_t w, uint32_t h, float *az, float *el) {
  const float TAU = 6.28318530717958647692f, FRAC_PI_2 = 1.57079632679489661923f;
  for (uint32_t i = 0; i < w; i++) {
    const float a = ((float)i + 0.5f) / (float)w * TAU;
    az[2 * i] = std::cos(a);
    az[2 * i + 1] = std::sin(a);
  }
  for (uint32_t j = 0; j < h; j++) {
    const float e = ((float)j + 0.5f) / (float)h * PI_F - FRAC_PI_2;
    el[2 * j] = std::cos(e);
    el[2 * j + 1] = std::sin(e);
  }
}

void lw_expf_tabulated(const float *x, size_t n, float *out) {
  for (size_t i = 0; i < n; i++) out[i] = expf_tabulated(x[i]);
}

void lw_camera_basis(float yaw_deg, float pitch_deg, float fwd[3], float right[3], float up[3]) {
  const float RADS = PI_F / 180.0f;  // f32::to_radians
  float y = yaw_deg * RADS, p = pitch_deg * RADS;
  float f[3] = {-std::sin(y) * std::cos(p), -std::sin(p), std::cos(y) * std::cos(p)};
  normalize3(f);
  // r = f x Y
  float r[3] = {f[1] * 0.0f - 1.0f * f[2], f[2] * 0.0f - 0.0f * f[0], f[0] * 1.0f - 0.0f * f[1]};
  if (r[0] * r[0] + r[1] * r[1] + r[2] * r[2] < 1e-8f) {
    r[0] = -std::cos(y);
    r[1] = 0.0f;
    r[2] = -std::sin(y);
  }
  normalize3(r);
  float u[3] = {r[1] * f[2] - f[1] * r[2], r[2] * f[0] - f[2] * r[0], r[0] * f[1] - f[0] * r[1]};
  normalize3(u);
  for (int k = 0; k < 3; k++) {
    fwd[k] = f[k];
    right[k] = r[k];
    up[k] = u[k];
  }
}

void lw_camera_rays_new(float yaw_deg, float pitch_deg, float fov_y_deg, uint32_t width,
                        uint32_t height, lw_camera_rays *out) {
  lw_camera_basis(yaw_deg, pitch_deg, out->forward, out->right, out->up);
  const float RADS = PI_F / 180.0f;  // f32::to_radians
  out->tan_half_fov = std::tan(fov_y_deg * RADS * 0.5f);
  out->aspect = (float)width / (float)height;
}

void lw_hq_params(const lw_hq_options *o, float emitter_intensity, uint32_t width, uint32_t height,
                  const lw_scene_info *hq, const float eye[3], const lw_camera_rays *rays,
                  const lw_hq_sky *s, double turbidity, lw_params *out) {
  lw_params p;
  std::memset(&p, 0, sizeof(p));
  auto f4 = [](float dst[4], const float v[3], float w) {
    dst[0] = v[0];
    dst[1] = v[1];
    dst[2] = v[2];
    dst[3] = w;
  };
  lw_preetham sky;
  lw_preetham_new(s->sun_dir, turbidity, &sky);
  float su[3], sv[3];
  lw_sun_texture_axes(s->sun_dir, su, sv);
  f4(p.cam_pos, eye, rays->tan_half_fov);
  f4(p.cam_fwd, rays->forward, rays->aspect);
  f4(p.cam_right, rays->right, 0.0f);
  f4(p.cam_up, rays->up, 0.0f);
  f4(p.sun_dir, s->sun_dir, std::cos(SUN_RADIUS));
  f4(p.sun_color, s->sun_color, SUN_RADIUS);
  f4(p.sky_sun, sky.sun, s->sky_light);
  f4(p.sky_zenith, sky.zenith, 0.0f);
  f4(p.perez_a, sky.perez[0], 0.0f);
  f4(p.perez_b, sky.perez[1], 0.0f);
  f4(p.perez_c, sky.perez[2], 0.0f);
  f4(p.perez_d, sky.perez[3], 0.0f);
  f4(p.perez_e, sky.perez[4], 0.0f);
  const float water[3] = {0.42f, 0.16f, 0.10f};
  f4(p.water_color, water, 1.333f);
  p.misc_f[0] = emitter_intensity;
  p.misc_f[1] = 1e-3f;
  p.misc_f[2] = 128.0f;  // sky.rs:590 SKY_CACHE_RESOLUTION
  p.misc_f[3] = 0.0f;
  p.grid_dims[0] = hq->dims[0];
  p.grid_dims[1] = hq->dims[1];
  p.grid_dims[2] = hq->dims[2];
  p.grid_dims[3] = hq->dims[0] * hq->dims[1] * hq->dims[2];
  for (int k = 0; k < 4; k++) p.off[k] = hq->offsets[k];
  p.off2[0] = hq->offsets[4];
  p.off2[1] = hq->atlas_size[0];
  p.off2[2] = hq->atlas_size[1];
  p.off2[3] = hq->quads_len;
  p.frame[0] = width;
  p.frame[1] = height;
  p.frame[2] = 0;
  p.frame[3] = o->max_bounces;
  p.misc_u[0] = o->seed;
  p.misc_u[1] = 1;
  p.misc_u[2] = hq->models_len;
  p.misc_u[3] = hq->emitter_count;
  p.see_through[0] = hq->see_through[0];
  p.see_through[1] = hq->see_through[1];
  p.emitter[0] = emitter_intensity > 0.0f ? o->emitter_sampling : LW_EMITTER_SAMPLING_NONE;
  p.emitter[1] = o->prevent_normal_emitter ? 1u : 0u;
  f4(p.sun_su, su, SUN_TEXTURE_BRIGHTNESS);
  f4(p.sun_sv, sv, 4.0f * SUN_RADIUS);
  if (hq->has_sun_sprite)
    for (int k = 0; k < 4; k++) p.sun_sprite[k] = (float)hq->sun_sprite[k];
  *out = p;
}

int32_t lw_world_height(const lw_world_spec *spec, int32_t x, int32_t z) {
  return terrain_height(spec->seed, x, z);
}

int lw_world_build(const lw_world_spec *spec, lw_world **out) {
  *out = nullptr;
  try {
    const uint32_t dx = spec->chunks_x, dy = spec->sections, dz = spec->chunks_z;
    if (dx == 0 || dy == 0 || dz == 0 || dx > 256 || dz > 256 || dy > 64)
      throw std::invalid_argument("world dims out of range");
    if (spec->cube_types == 0 || spec->cube_types + spec->cutout_types > 1000)
      throw std::invalid_argument("block type counts out of range");
    const int32_t gx = (int32_t)dx * 16, gy = (int32_t)dy * 16, gz = (int32_t)dz * 16;
    const uint64_t seed = spec->seed, tseed = spec->texture_seed;
    std::vector<int32_t> height((size_t)gx * gz);
    for (int32_t x = 0; x < gx; x++)
      for (int32_t z = 0; z < gz; z++) height[(size_t)x * gz + z] = terrain_height(seed, x, z);
    auto H = [&](int32_t x, int32_t z) { return height[(size_t)x * gz + z]; };
    auto in_grid = [&](int32_t x, int32_t y, int32_t z) {
      return x >= 0 && y >= 0 && z >= 0 && x < gx && y < gy && z < gz;
    };
    auto solid = [&](int32_t x, int32_t y, int32_t z) { return in_grid(x, y, z) && y <= H(x, z); };
    const int32_t wl = spec->water_level;
    auto water = [&](int32_t x, int32_t y, int32_t z) { return wl >= 0 && !solid(x, y, z) && y < wl; };
    const uint32_t n_cube = spec->cube_types, n_cut = spec->cutout_types;
    // Model ids: 0 air, 1..n_cube cubes, then the plants, then the water surface, then one
    // model per emitter (its own emittances).
    const uint32_t water_model = 1 + n_cube + n_cut;
    auto plant = [&](int32_t x, int32_t y, int32_t z) -> uint32_t {
      if (n_cut == 0) return 0;
      const int32_t h = H(x, z);
      if (y != h + 1 || y >= gy || (wl >= 0 && h + 1 < wl)) return 0;
      if (u01(hash5(seed, H_PLANT, (uint64_t)x, (uint64_t)z, 0)) >= spec->cutout_fraction) return 0;
      return 1 + n_cube + (uint32_t)(hash5(seed, H_PLANT_TYPE, (uint64_t)x, (uint64_t)z, 0) % n_cut);
    };
    auto exposed = [&](int32_t x, int32_t y, int32_t z) {
      static const int32_t N[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
      for (auto &n : N) {
        int32_t a = x + n[0], b = y + n[1], c = z + n[2];
        if (in_grid(a, b, c) && !solid(a, b, c)) return true;
      }
      return false;
    };

    // Atlas: one 16 x 16 sprite per cube type, plant type and the water, 32 to a row.
    const uint32_t n_sprites = n_cube + n_cut + 1;
    const uint32_t aw = 16u * std::min(n_sprites, 32u), ah = 16u * ((n_sprites + 31) / 32);
    std::vector<uint32_t> atlas((size_t)aw * ah, 0u);
    auto sprite_xy = [&](uint32_t k, uint32_t &sx, uint32_t &sy) {
      sx = (k % 32) * 16;
      sy = (k / 32) * 16;
    };
    auto enc = [](double a) {  // linear albedo to its sRGB-ish byte (a^(1/2.2))
      return (uint32_t)std::lround(255.0 * std::pow(a, 1.0 / 2.2));
    };
    for (uint32_t k = 0; k < n_sprites; k++) {
      uint32_t sx, sy;
      sprite_xy(k, sx, sy);
      const bool is_plant = k >= n_cube && k < n_cube + n_cut;
      const bool is_water = k == n_cube + n_cut;
      for (uint32_t ty = 0; ty < 16; ty++)
        for (uint32_t tx = 0; tx < 16; tx++) {
          uint32_t rgb[3];
          for (int c = 0; c < 3; c++) {
            double a = 0.1 + 0.8 * u01(hash5(tseed, H_TEXEL, k, ty * 16 + tx, (uint64_t)c));
            if (is_water) a = (c == 2 ? 0.6 : (c == 1 ? 0.35 : 0.15)) + 0.05 * (a - 0.5);
            rgb[c] = enc(a);
          }
          uint32_t alpha = 255;
          if (is_plant) alpha = u01(hash5(tseed, H_ALPHA, k, ty * 16 + tx, 0)) < 0.5 ? 255u : 0u;
          atlas[(size_t)(sy + ty) * aw + sx + tx] = rgb[0] | (rgb[1] << 8) | (rgb[2] << 16) | (alpha << 24);
        }
    }

    std::vector<lw_model> models(1, lw_model{0, 0, 0, 0});
    std::vector<lw_quad> quads;
    std::vector<uint32_t> quad_type, quad_emitter;  // kept level with `quads`
    auto tag = [&](uint32_t type, uint32_t emitter) {
      quad_type.resize(quads.size(), type);
      quad_emitter.resize(quads.size(), emitter);
    };
    const float no_emit[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t t = 0; t < n_cube; t++) {
      uint32_t sx, sy;
      sprite_xy(t, sx, sy);
      models.push_back(lw_model{(uint32_t)quads.size(), 6, 0, 0});
      cube_faces(quads, LW_ALPHA_OPAQUE | LW_FLAG_FRONT_ONLY, sx, sy, no_emit);
      tag(t, LW_NO_GROUP);
    }
    for (uint32_t t = 0; t < n_cut; t++) {
      uint32_t sx, sy;
      sprite_xy(n_cube + t, sx, sy);
      models.push_back(lw_model{(uint32_t)quads.size(), 4, 0, 0});
      cross_faces(quads, sx, sy);
      tag(n_cube + t, LW_NO_GROUP);
    }
    {
      uint32_t sx, sy;
      sprite_xy(n_cube + n_cut, sx, sy);
      models.push_back(lw_model{(uint32_t)quads.size(), 1, 0, 0});
      const float p0[3] = {0.0f, 0.875f, 0.0f}, eu[3] = {1, 0, 0}, ev[3] = {0, 0, 1};
      quads.push_back(make_quad(p0, eu, ev, LW_FLAG_WATER, 0.0f, sx, sy));
      tag(LW_NO_GROUP, LW_NO_GROUP);
    }
    // Seeded emitter columns: their surface block becomes an emitter.
    std::vector<uint8_t> emitter_column((size_t)gx * gz, 0);
    if (spec->emitter_blocks > 0) {
      if (gx <= 16 || gz <= 16) throw std::invalid_argument("world too small for emitter blocks");
      const uint32_t want = std::min<uint32_t>(spec->emitter_blocks, (uint32_t)((gx - 16) * (gz / 2 - 8)));
      uint32_t placed = 0;
      for (uint64_t k = 0; placed < want; k++) {
        const int32_t x = 8 + (int32_t)(hash5(seed, H_EMIT_COLUMN, k, 0, 0) % (uint64_t)(gx - 16));
        const int32_t z = gz / 2 + (int32_t)(hash5(seed, H_EMIT_COLUMN, k, 1, 0) % (uint64_t)(gz / 2 - 8));
        uint8_t &c = emitter_column[(size_t)x * gz + z];
        if (!c && H(x, z) < gy) {
          c = 1;
          placed++;
        }
      }
    }

    const uint64_t nb = (uint64_t)dx * dy * dz;
    auto brick_xyz = [&](uint64_t i, int32_t &bx, int32_t &by, int32_t &bz) {
      bx = (int32_t)(i % dx);
      by = (int32_t)(i / (dx * dz));
      bz = (int32_t)((i / dx) % dz);
    };
    lw_world *w = (lw_world *)std::calloc(1, sizeof(lw_world));
    if (!w) throw std::bad_alloc();
    std::unique_ptr<lw_world, void (*)(lw_world *)> guard(w, lw_world_free);

    // The interned volume: per brick whether every block occludes (voxel.rs:119-130), and
    // every voxel's model id; an emitter block gets a model of its own (its emittances).
    std::vector<uint8_t> all_occluding(nb, 0);
    for (uint64_t b = 0; b < nb; b++) {
      int32_t bx, by, bz;
      brick_xyz(b, bx, by, bz);
      bool all = true;
      for (int32_t x = 0; x < 16 && all; x++)
        for (int32_t z = 0; z < 16 && all; z++)
          if (by * 16 + 15 > H(bx * 16 + x, bz * 16 + z)) all = false;
      all_occluding[b] = all ? 1 : 0;
    }
    std::vector<uint16_t> vox(nb * 4096);
    std::vector<uint8_t> flags(models.size(), 0);
    for (uint32_t m = 1; m <= n_cube; m++) flags[m] = LW_MODEL_OCCLUDES;
    uint32_t n_emitters = 0;
    for (uint64_t b = 0; b < nb; b++) {
      int32_t bx, by, bz;
      brick_xyz(b, bx, by, bz);
      for (uint32_t local = 0; local < 4096; local++) {
        const int32_t lx = local & 15, lz = (local >> 4) & 15, ly = local >> 8;
        const int32_t x = bx * 16 + lx, y = by * 16 + ly, z = bz * 16 + lz;
        uint32_t id = 0;
        if (solid(x, y, z)) {
          w->solid_cells++;
          const bool column_emitter = emitter_column[(size_t)x * gz + z] && y == H(x, z);
          if (column_emitter ||
              (spec->emitter_fraction > 0.0f && exposed(x, y, z) &&
               u01(hash5(seed, H_EMIT, (uint64_t)x, (uint64_t)y, (uint64_t)z)) < spec->emitter_fraction)) {
            const uint32_t t = (uint32_t)(hash5(seed, H_TYPE, (uint64_t)x, (uint64_t)y, (uint64_t)z) % n_cube);
            uint32_t sx, sy;
            sprite_xy(t, sx, sy);
            float em[6];
            const uint32_t e_index = n_emitters++;
            for (int f = 0; f < 6; f++)
              em[f] = column_emitter ? 1.0f
                                     : (float)(1.0 + 3.0 * u01(hash5(tseed, H_EMITTANCE, e_index, (uint64_t)f, 0)));
            id = (uint32_t)models.size();
            models.push_back(lw_model{(uint32_t)quads.size(), 6, 0, 0});
            flags.push_back(LW_MODEL_OCCLUDES | LW_MODEL_EMITS);
            cube_faces(quads, LW_ALPHA_OPAQUE | LW_FLAG_FRONT_ONLY, sx, sy, em);
            tag(t, e_index);
          } else {
            id = 1 + (uint32_t)(hash5(seed, H_TYPE, (uint64_t)x, (uint64_t)y, (uint64_t)z) % n_cube);
          }
        } else if (water(x, y, z)) {
          w->water_cells++;
          if (y == wl - 1 && !solid(x, y + 1, z)) id = water_model;
        } else {
          id = plant(x, y, z);
          if (id) w->cutout_cells++;
        }
        if (id > 0xFFFF) throw std::runtime_error("more than 65535 distinct models");
        vox[b * 4096 + local] = (uint16_t)id;
      }
    }

    // The layout (voxel.rs:305-342): EMPTY, buried, UNIFORM or stored bricks, and the emitter
    // sites in brick then voxel order.
    auto brick_index = [&](int32_t bx, int32_t by, int32_t bz) -> int64_t {
      if (bx < 0 || by < 0 || bz < 0 || bx >= (int32_t)dx || by >= (int32_t)dy || bz >= (int32_t)dz) return -1;
      return ((int64_t)by * dz + bz) * dx + bx;
    };
    std::vector<uint32_t> top(nb, LW_EMPTY);
    std::vector<uint32_t> bricks;
    std::vector<std::pair<std::array<int32_t, 3>, uint32_t>> sites;
    for (uint64_t b = 0; b < nb; b++) {
      int32_t bx, by, bz;
      brick_xyz(b, bx, by, bz);
      const uint16_t *ids = vox.data() + b * 4096;
      bool any = false;
      for (uint32_t local = 0; local < 4096; local++) {
        any = any || ids[local] != 0;
        if (flags[ids[local]] & LW_MODEL_EMITS)
          sites.push_back({{bx * 16 + (int32_t)(local & 15), by * 16 + (int32_t)(local >> 8),
                            bz * 16 + (int32_t)((local >> 4) & 15)},
                           ids[local]});
      }
      if (!any) continue;  // an empty section stays EMPTY
      bool buried = all_occluding[b] != 0;
      static const int32_t N[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
      for (auto &n : N) {
        int64_t nbi = brick_index(bx + n[0], by + n[1], bz + n[2]);
        if (nbi < 0 || !all_occluding[nbi]) buried = false;
      }
      if (buried) {
        w->buried_bricks++;
        uint32_t first = 1;
        for (uint32_t k = 0; k < 4096; k++)
          if (ids[k] != 0) {
            first = ids[k];
            break;
          }
        top[b] = LW_UNIFORM | first;
        continue;
      }
      bool uniform = true;
      for (uint32_t k = 1; k < 4096 && uniform; k++) uniform = ids[k] == ids[0];
      if (uniform) {
        w->uniform_bricks++;
        top[b] = LW_UNIFORM | ids[0];
        continue;
      }
      top[b] = (uint32_t)(bricks.size() / LW_BRICK_WORDS);
      for (uint32_t k = 0; k < 4096; k += 2) bricks.push_back((uint32_t)ids[k] | ((uint32_t)ids[k + 1] << 16));
      w->bricks++;
    }

    std::vector<lw_emitter> emitters;
    std::vector<uint32_t> egrid;
    emitter_list(models, sites, emitters, egrid);

    std::vector<uint32_t> grid;
    grid.reserve(nb * 2 + bricks.size() + 1);
    const uint32_t off_top = 0;
    grid.insert(grid.end(), top.begin(), top.end());
    const uint32_t off_bricks = (uint32_t)grid.size();
    grid.insert(grid.end(), bricks.begin(), bricks.end());
    const uint32_t off_ov_top = (uint32_t)grid.size();
    grid.insert(grid.end(), nb, LW_EMPTY);
    const uint32_t off_ov_cells = (uint32_t)grid.size();
    const uint32_t off_ov_quads = (uint32_t)grid.size();
    grid.push_back(0u);

    w->grid = to_heap(grid);
    w->grid_len = grid.size();
    w->models = to_heap(models);
    w->models_len = models.size();
    w->quads = to_heap(quads);
    w->quads_len = quads.size();
    w->atlas = to_heap(atlas);
    w->atlas_len = atlas.size();
    w->emitters = to_heap(emitters);
    w->emitters_len = emitters.size();
    w->emitter_grid = to_heap(egrid);
    w->emitter_grid_len = egrid.size();
    w->quad_type = to_heap(quad_type);
    w->quad_emitter = to_heap(quad_emitter);
    w->voxels = to_heap(vox);
    w->voxels_len = vox.size();
    w->brick_occludes = to_heap(all_occluding);
    w->model_flags = to_heap(flags);
    lw_scene_info &info = w->info;
    std::memset(&info, 0, sizeof(info));
    info.dims[0] = dx;
    info.dims[1] = dy;
    info.dims[2] = dz;
    const uint32_t offs[5] = {off_top, off_bricks, off_ov_top, off_ov_cells, off_ov_quads};
    std::memcpy(info.offsets, offs, sizeof(offs));
    info.atlas_size[0] = aw;
    info.atlas_size[1] = ah;
    info.quads_len = (uint32_t)quads.size();
    info.models_len = (uint32_t)models.size();
    info.emitter_count = (uint32_t)sites.size();
    info.see_through[0] = (uint32_t)quads.size();
    info.see_through[1] = 0;
    *out = guard.release();
    return 0;
  } catch (const std::exception &e) {
    g_host_error = e.what();
    return LW_EINVAL;
  }
}

void lw_world_free(lw_world *w) {
  if (!w) return;
  std::free(w->grid);
  std::free(w->models);
  std::free(w->quads);
  std::free(w->atlas);
  std::free(w->emitters);
  std::free(w->emitter_grid);
  std::free(w->quad_type);
  std::free(w->quad_emitter);
  std::free(w->voxels);
  std::free(w->brick_occludes);
  std::free(w->model_flags);
  std::free(w);
}

long lw_dusk_sky_maps(const float sun_dir[3], uint32_t lw, uint32_t lh, uint32_t gw, uint32_t gh,
                      float *out, size_t cap) {
  const size_t need = ((size_t)lw * lh + (size_t)gw * gh) * 3;
  if (cap < need) return -1;
  const double sun_az = std::atan2((double)sun_dir[2], (double)sun_dir[0]);
  const double sun_alt = std::asin(std::max(-1.0, std::min(1.0, (double)sun_dir[1])));
  // Linear radiance of a dusk sky towards azimuth az (radians from +x towards +z) and
  // elevation el: a deep blue zenith, a warm band along the horizon brightest under the
  // sun, and a dark ground below.
  auto sky = [&](double az, double el, double out3[3]) {
    const double zen[3] = {0.015, 0.025, 0.070};
    const double hor[3] = {0.180, 0.110, 0.080};
    const double glow[3] = {0.900, 0.380, 0.110};
    const double up = std::max(0.0, std::sin(el));
    const double band = std::pow(1.0 - up, 4.0);
    const double dphi = std::remainder(az - sun_az, 2.0 * M_PI);
    const double toward = 0.5 + 0.5 * std::cos(dphi);
    const double g = std::exp(-std::fabs(el - sun_alt) / 0.12) * toward * toward;
    for (int c = 0; c < 3; c++) {
      double v = zen[c] + (hor[c] - zen[c]) * band + glow[c] * g;
      if (el < 0.0) v *= 0.25 + 0.75 * std::exp(el / 0.05);
      out3[c] = v;
    }
  };
  size_t k = 0;
  for (int m = 0; m < 2; m++) {
    const uint32_t W = m == 0 ? lw : gw, Hh = m == 0 ? lh : gh;
    for (uint32_t j = 0; j < Hh; j++) {
      const double el = ((double)j + 0.5) / Hh * M_PI - M_PI_2;
      for (uint32_t i = 0; i < W; i++) {
        const double az = ((double)i + 0.5) / W * 2.0 * M_PI;
        double c[3];
        sky(az, el, c);
        for (int ch = 0; ch < 3; ch++) out[k++] = (float)c[ch];
      }
    }
  }
  return (long)need;
}

}  // extern "C"

// ---- post-processing (render/post.rs) ----

namespace {

struct Hable {
  float a, b, c, d, e, f, w;
  // post.rs:70-83: 1 / h(W) in float
  float white_scale() const {
    const float num = w * (a * w + c * b) + d * e;
    const float den = w * (a * w + b) + d * f;
    return 1.0f / (num / den - e / f);
  }
  // post.rs:113-121: A k^2, B k, W / k through f64
  Hable brightened(double k) const {
    Hable r = *this;
    r.a = (float)((double)a * k * k);
    r.b = (float)((double)b * k);
    r.w = (float)((double)w / k);
    return r;
  }
  // post.rs:122-131: the positive root of the quadratic in y = 2 k m
  double mid_grey_gain(double m) const {
    const double aa = a, bb = b, cb = (double)(c * b), de = (double)(d * e), df = (double)(d * f),
                 ef = (double)(e / f);
    const double q = m / (double)white_scale() + ef;
    const double qa = aa * (1.0 - q), qb = cb - q * bb, qc = de - q * df;
    const double y = (-qb + std::sqrt(qb * qb - 4.0 * qa * qc)) / (2.0 * qa);
    return y / (2.0 * m);
  }
};

constexpr Hable kFilmicWorlds = {0.15f, 0.5f, 0.1f, 0.2f, 0.02f, 0.3f, 11.2f};  // post.rs:55-63

uint32_t display_byte(float v) {  // post.rs:141-144,165 with Rust's saturating `as u32`
  float x = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
  const float d = lw::display_pow(x) * 255.0f + 0.5f;
  if (!(d > 0.0f)) return 0;
  return (uint32_t)d;
}

}  // namespace

float lw::display_pow(float x) { return std::pow(x, 1.0f / 2.2f); }
float lw::libm_powf(float x, float y) { return std::pow(x, y); }
float lw::libm_log2f(float x) { return std::log2(x); }

void lw_hable_curve(uint32_t op, float out[8]) {
  Hable h = kFilmicWorlds;
  if (op == LW_TONEMAP_FILMIC_MID_GREY) h = kFilmicWorlds.brightened(kFilmicWorlds.mid_grey_gain(0.18));
  const float v[8] = {h.a, h.b, h.c, h.d, h.e, h.f, h.w, h.white_scale()};
  std::memcpy(out, v, sizeof(v));
}

void lw_encode_thresholds(float out[255]) {
  // display_byte is non-decreasing over [0, 1] (tests check every f32 there), so each
  // threshold is found by bisection over the bit patterns of [0, 1].
  for (uint32_t k = 1; k <= 255; k++) {
    uint32_t lo = 0, hi = 0x3f800000u;  // display_byte(1) = 255 >= k
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      float v;
      std::memcpy(&v, &mid, 4);
      if (display_byte(v) >= k)
        hi = mid;
      else
        lo = mid + 1;
    }
    std::memcpy(&out[k - 1], &lo, 4);
  }
}

int lw_write_pfm(const char *path, const float *hdr, uint32_t w, uint32_t h) {
  FILE *f = path ? std::fopen(path, "wb") : nullptr;
  if (!f) {
    g_host_error = std::string("cannot open ") + (path ? path : "(null)");
    return -1;
  }
  const std::string head = "PF\n" + std::to_string(w) + " " + std::to_string(h) + "\n-1.0\n";
  bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
  for (uint32_t yy = 0; ok && yy < h; yy++) {
    const float *row = hdr + (size_t)(h - 1 - yy) * w * 3;
    ok = std::fwrite(row, 4, (size_t)w * 3, f) == (size_t)w * 3;  // little-endian host
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) g_host_error = std::string("writing ") + path;
  return ok ? 0 : -1;
}
