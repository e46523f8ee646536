// ORACLE -- test infrastructure only. Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this code, and only as the checker or the
// timed CPU baseline; the product path never calls it.
//
// An op-for-op C++ restatement of the reference's path-tracing kernel crate
// (crates/lumenwright-kernel/src/{rng,math,scene,trace,integrator}.rs) and of its CPU
// backend (crates/lumenwright/src/render/cpu/mod.rs:39-87). Every f32 operation is
// written in the reference's evaluation order; the build uses -ffp-contract=off and no
// fast-math, so on x86-64 (SSE) each operation rounds exactly as the reference's Rust
// does. Deviations from "what the reference would compute on Linux":
//   * none in the arithmetic: exp is glibc's expf, which Rust std's f32::exp calls;
//   * an out-of-bounds slice index, a panic in Rust, throws OracleError here.
//
// This file is included three times: by lw_oracle.cpp (LWO_COUNTING 0), by
// lw_oracle_counted.cpp (LWO_COUNTING 1), which adds the algorithmic-byte counters of the
// measurement rules (8 B per DDA step, 80 B per quad tested, ...), and by lw_oracle_f64.cpp
// (LWO_F64 1), where `real` is double and the transcendental functions are libm's double
// ones: the f64 render against which the f32 path's rounding error is stated.
//
// Parity pins (tests/test_oracle_*.py): the reference's own known-answer tests in
// crates/lumenwright-kernel/src/tests.rs:128-831 (exact hits, emission 6.5, sky maps,
// lamp quadrature, RNG KAT, math bounds), the f64 Chunky transcriptions of
// crates/lumenwright/src/render/sky.rs:980-1397, and brute-force f64 ray/box and
// ray/parallelogram checks.

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace LWO_NS {

struct OracleError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ---- layout: crates/lumenwright-kernel/src/scene.rs:44-228 ----
struct Quad {
  float p0[4], eu[4], ev[4], uv[4], uvt[4];
};
struct Model {
  uint32_t quad_offset, quad_count, flags, pad;
};
struct Emitter {
  float pos[4];
  uint32_t faces[4];
};
struct Params {
  float cam_pos[4], cam_fwd[4], cam_right[4], cam_up[4], sun_dir[4], sun_color[4], sky_sun[4],
      sky_zenith[4], perez_a[4], perez_b[4], perez_c[4], perez_d[4], perez_e[4], water_color[4],
      misc_f[4];
  uint32_t grid_dims[4], off[4], off2[4], frame[4], misc_u[4], see_through[4], emitter[4];
  float sun_su[4], sun_sv[4], sun_sprite[4], sky_map[4], water_plane[4];
};
static_assert(sizeof(Quad) == 80, "GpuQuad is 80 bytes (trace.rs:329)");
static_assert(sizeof(Params) == 432, "Params is 432 bytes (trace.rs:330)");
static_assert(sizeof(Model) == 16, "GpuModel is 16 bytes (trace.rs:331)");
static_assert(sizeof(Emitter) == 32, "GpuEmitter is 32 bytes (trace.rs:332)");

// SceneView with slice lengths (scene.rs:231-259).
struct View {
  const Params *params;
  const uint32_t *grid;
  size_t grid_len;
  const Model *models;
  size_t models_len;
  const Quad *quads;
  size_t quads_len;
  const uint32_t *atlas;
  size_t atlas_len;
  const Emitter *emitters;
  size_t emitters_len;
  const uint32_t *emitter_grid;
  size_t emitter_grid_len;
  const float *sky;
  size_t sky_len;
};

// Hit (trace.rs:11-22) plus the final DDA cell; layout shared with lw_hit.
struct HitOut {
  float t;
  uint32_t quad, texel;
  float s, tt;
  int32_t cell[3];
  uint32_t hit, pad;
};

// Algorithmic scene-buffer bytes and trace calls (the measurement rules), and how often the
// rarer branches of render_sample and shadow were taken (so that a test can show a scene
// reaches them).
struct Counters {
  uint64_t bytes;
  uint64_t dda_steps;
  uint64_t quads_tested;
  uint64_t traces;
  uint64_t samples;
  uint64_t sky_lookups;
  uint64_t see_through_drawn;   // a see-through quad drawn instead of the surface met
  uint64_t see_through_forced;  // ... of which translucent, taken by its alpha draw
  uint64_t translucent_passes;  // a path passing through a translucent surface
  uint64_t alpha_zero_rejects;  // a translucent quad met at a texel of alpha 0
  uint64_t shadow_water_exits;  // a shadow segment leaving water (Beer-Lambert)
  uint64_t shadow_translucent;  // a shadow segment filtered by a translucent surface
  uint64_t water_interfaces;    // a path event at a water quad
  uint64_t pad;
};

#if LWO_COUNTING
static thread_local Counters *tl_counters = nullptr;
#define LWO_COUNT(field, amount)                   \
  do {                                             \
    if (tl_counters) tl_counters->field += (amount); \
  } while (0)
#else
#define LWO_COUNT(field, amount) \
  do {                           \
  } while (0)
#endif

#if LWO_COUNTING
// The device's walk stops once every brick left toward the ray's octant is EMPTY with no
// overlay (no quad can be tested after that, so its result is the same "no hit"), and the
// counted walk stops there too, so that the counts are those of the walk the device runs.
// Per brick, bit k (k = [+x] | [+y] << 1 | [+z] << 2, step 0 counted as +) is set when the
// box from the brick to the grid's end toward octant k holds a brick that is not EMPTY.
static thread_local const std::vector<uint8_t> *tl_octant_boxes = nullptr;
#endif

template <typename T>
static inline const T &at(const T *buf, size_t len, size_t i, const char *what) {
  if (i >= len) throw OracleError(std::string("index out of bounds in ") + what);
  return buf[i];
}

// ---- rng.rs:5-34 ----
static inline uint32_t pcg_hash(uint32_t v) {
  uint32_t state = v * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
  return (word >> 22) ^ word;
}
// The arithmetic type: float as the reference computes, or double for the f64 render that
// states the f32 path's own rounding error (LWO_F64).
#ifndef LWO_F64
#define LWO_F64 0
#endif
#if LWO_F64
typedef double real;
#else
typedef float real;
#endif

struct Rng {
  uint32_t state;
  static Rng make(uint32_t pixel, uint32_t sample, uint32_t seed) {
    return Rng{pcg_hash(pixel ^ pcg_hash(sample + pcg_hash(seed)))};
  }
  real next() {
    state = pcg_hash(state);
    return (real)(state >> 8) * (real)(1.0f / 16777216.0f);
  }
};

// ---- Rust `as` casts: truncate toward zero, saturate, NaN -> 0 ----
static inline int32_t as_i32(real x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return INT32_MAX;
  if (x < -2147483648.0f) return INT32_MIN;
  return (int32_t)x;
}
static inline uint32_t as_u32(real x) {
  if (x != x) return 0;
  if (x >= 4294967296.0f) return UINT32_MAX;
  if (x <= 0.0f) return 0;
  return (uint32_t)x;
}
static inline float bits_f(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
static inline uint32_t f_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// ---- math.rs ----
#if LWO_F64
// The f64 render: every operation in double, the transcendental functions libm's double
// ones in place of the reference's f32 approximations; the constants are the reference's.
static const real PI_F = 3.14159265358979323846;
static const real FRAC_PI_2_F = 1.57079632679489661923;
static const real SQRT_2_F = 1.41421356237309504880;
static const real LN_2_F = 0.693147180559945309417;
static const real GAMMA = 2.2f;

static inline real f_sqrt(real x) { return std::sqrt(x); }
static inline real f_floor(real x) { return std::floor(x); }
static inline real f_exp(real x) { return std::exp(x); }
static inline real f_abs(real x) { return x < 0.0 ? -x : x; }
static inline real f_min(real a, real b) { return a < b ? a : b; }
static inline real f_max(real a, real b) { return a > b ? a : b; }
static inline real f_clamp(real x, real lo, real hi) { return f_min(f_max(x, lo), hi); }
static inline real f_rem(real a, real b) { return std::fmod(a, b); }
static inline void sin_cos_turns(real u, real &s_out, real &c_out) {
  const real a = u * (2.0 * 3.14159265358979323846);
  s_out = std::sin(a);
  c_out = std::cos(a);
}
static inline real f_acos(real x) { return std::acos(f_clamp(x, -1.0, 1.0)); }
static inline real f_atan2(real y, real x) { return (x == 0.0 && y == 0.0) ? 0.0 : std::atan2(y, x); }
static inline real f_ln(real x) { return x <= 0.0 ? (x == 0.0 ? -INFINITY : NAN) : std::log(x); }
static inline real gamma_to_linear(real c) { return c <= 0.0 ? 0.0 : f_exp(GAMMA * f_ln(c)); }
#else
static const float PI_F = 3.14159265358979323846f;
static const float FRAC_PI_2_F = 1.57079632679489661923f;
static const float SQRT_2_F = 1.41421356237309504880f;
static const float LN_2_F = 0.693147180559945309417f;
static const float GAMMA = 2.2f;

static inline float f_sqrt(float x) { return std::sqrt(x); }
static inline float f_floor(float x) { return std::floor(x); }
static inline float f_exp(float x) { return ::expf(x); }  // Rust std f32::exp -> libm expf
static inline float f_abs(float x) { return x < 0.0f ? -x : x; }
static inline float f_min(float a, float b) { return a < b ? a : b; }
static inline float f_max(float a, float b) { return a > b ? a : b; }
static inline float f_clamp(float x, float lo, float hi) { return f_min(f_max(x, lo), hi); }
static inline float f_rem(float a, float b) { return std::fmod(a, b); }  // Rust f32 `%`

// math.rs:44-64
static inline void sin_cos_turns(float u, float &s_out, float &c_out) {
  const float ROUND = 12582912.0f;
  float t = u * 4.0f;
  float q = (t + ROUND) - ROUND;
  float x = (t - q) * FRAC_PI_2_F;
  float x2 = x * x;
  float s = x * (1.0f +
                 x2 * (-1.0f / 6.0f +
                       x2 * (1.0f / 120.0f + x2 * (-1.0f / 5040.0f + x2 * (1.0f / 362880.0f)))));
  float c = 1.0f +
            x2 * (-0.5f +
                  x2 * (1.0f / 24.0f +
                        x2 * (-1.0f / 720.0f +
                              x2 * (1.0f / 40320.0f + x2 * (-1.0f / 3628800.0f)))));
  int32_t k = as_i32(q) & 3;
  bool odd = (k & 1) != 0;
  float s2 = odd ? c : s;
  float c2 = odd ? -s : c;
  if (k & 2) {
    s_out = -s2;
    c_out = -c2;
  } else {
    s_out = s2;
    c_out = c2;
  }
}

// math.rs:86-103 (Abramowitz & Stegun 4.4.46)
static inline float f_acos(float x) {
  float a = f_clamp(f_abs(x), 0.0f, 1.0f);
  float p = 1.5707963f +
            a * (-0.2145988f +
                 a * (0.08897899f +
                      a * (-0.0501743f +
                           a * (0.03089188f +
                                a * (-0.017088126f + a * (0.00667009f + a * -0.0012624911f))))));
  float r = f_sqrt(1.0f - a) * p;
  return x < 0.0f ? PI_F - r : r;
}

// math.rs:113-142 (Abramowitz & Stegun 4.4.49)
static inline float f_atan2(float y, float x) {
  float ax = f_abs(x), ay = f_abs(y);
  float lo = f_min(ax, ay), hi = f_max(ax, ay);
  if (hi == 0.0f) return 0.0f;
  float t = lo / hi;
  float t2 = t * t;
  float p =
      t * (1.0f +
           t2 * (-0.33333145f +
                 t2 * (0.19993551f +
                       t2 * (-0.142089f +
                             t2 * (0.10656264f +
                                   t2 * (-0.07528964f +
                                         t2 * (0.042909614f +
                                               t2 * (-0.016165737f + t2 * 0.0028662257f))))))));
  float r = ay > ax ? FRAC_PI_2_F - p : p;
  r = x < 0.0f ? PI_F - r : r;
  return y < 0.0f ? -r : r;
}

// math.rs:182-203
static inline float f_ln(float x) {
  if (x <= 0.0f) return x == 0.0f ? -INFINITY : NAN;
  uint32_t bits = f_bits(x);
  int32_t e = (int32_t)((bits >> 23) & 0xFFu) - 127;
  float m = bits_f((bits & 0x007FFFFFu) | 0x3F800000u);
  if (m > SQRT_2_F) {
    m *= 0.5f;
    e += 1;
  }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float series =
      2.0f * s *
      (1.0f + s2 * (1.0f / 3.0f + s2 * (1.0f / 5.0f + s2 * (1.0f / 7.0f + s2 * (1.0f / 9.0f)))));
  return (float)e * LN_2_F + series;
}

// math.rs:213-216
static inline float gamma_to_linear(float c) { return c <= 0.0f ? 0.0f : f_exp(GAMMA * f_ln(c)); }

#endif  // LWO_F64
// math.rs:218-344
struct V3 {
  real x, y, z;
};
static inline V3 v3(real x, real y, real z) { return V3{x, y, z}; }
static inline V3 from4(const float a[4]) { return V3{a[0], a[1], a[2]}; }
static inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator*(V3 a, real s) { return V3{a.x * s, a.y * s, a.z * s}; }
static inline V3 operator/(V3 a, real s) { return a * (1.0f / s); }
static inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
static inline real dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline real length(V3 a) { return f_sqrt(dot(a, a)); }
static inline V3 normalize(V3 a) {
  real l = length(a);
  return l > 0.0f ? a * (1.0f / l) : a;
}
static inline V3 mul(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline real max_elem(V3 a) { return f_max(a.x, f_max(a.y, a.z)); }
static inline real get(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline V3 transmittance(V3 a, real t) {
  return V3{f_exp(-a.x * t), f_exp(-a.y * t), f_exp(-a.z * t)};
}
static inline V3 lerp(V3 a, V3 b, real t) { return a + (b - a) * t; }
static const V3 ZERO3 = {0.0f, 0.0f, 0.0f};
static const V3 ONE3 = {1.0f, 1.0f, 1.0f};

// math.rs:347-356 (Duff et al. 2017)
static inline void onb(V3 n, V3 &t_out, V3 &b_out) {
  real sign = n.z >= 0.0f ? 1.0f : -1.0f;
  real a = -1.0f / (sign + n.z);
  real b = n.x * n.y * a;
  t_out = V3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
  b_out = V3{b, sign + n.y * n.y * a, -n.y};
}

// ---- scene.rs ----
static const uint32_t EMPTY = 0xFFFFFFFFu;
static const uint32_t UNIFORM = 0x80000000u;
static const uint32_t BRICK_WORDS = 2048u;
static const uint32_t OV_CELL_WORDS = 8192u;
static const uint32_t ALPHA_CUTOUT = 1u, ALPHA_TRANSLUCENT = 2u, FLAG_ALPHA_MASK = 3u;
static const uint32_t FLAG_WATER = 1u << 2, FLAG_FRONT_ONLY = 1u << 4;
static const uint32_t EMITTER_SAMPLING_NONE = 0, EMITTER_SAMPLING_ONE = 1,
                      EMITTER_SAMPLING_ONE_BLOCK = 2, EMITTER_SAMPLING_ALL = 3;
static const uint32_t EMITTER_GRID_HEADER = 8u;

static inline uint32_t q_flags(const Quad &q) { return f_bits(q.p0[3]); }
static inline uint32_t q_tint(const Quad &q) { return f_bits(q.eu[3]); }
static inline real q_emittance(const Quad &q) { return q.ev[3]; }
static inline V3 q_normal(const Quad &q) { return cross(from4(q.ev), from4(q.eu)); }

static inline uint32_t grid_at(const View &s, size_t i) {
  return at(s.grid, s.grid_len, i, "grid");
}
static inline const Quad &quad_at(const View &s, size_t i) {
  return at(s.quads, s.quads_len, i, "quads");
}

// scene.rs:264-267
static inline V3 grid_size(const View &s) {
  const uint32_t *d = s.params->grid_dims;
  return V3{(real)(d[0] * 16u), (real)(d[1] * 16u), (real)(d[2] * 16u)};
}

// scene.rs:288-294
static inline uint32_t texel(const View &s, real u, real v) {
  int32_t w = (int32_t)s.params->off2[1];
  int32_t h = (int32_t)s.params->off2[2];
  int32_t x = as_i32(f_floor(u));
  int32_t y = as_i32(f_floor(v));
  if (w - 1 < 0 || h - 1 < 0) throw OracleError("empty atlas");
  x = x < 0 ? 0 : (x > w - 1 ? w - 1 : x);
  y = y < 0 ? 0 : (y > h - 1 ? h - 1 : y);
  LWO_COUNT(bytes, 4);
  return at(s.atlas, s.atlas_len, (size_t)(uint32_t)(y * w + x), "atlas");
}

// scene.rs:299-306
static inline void unpack_rgba(uint32_t c, real out[4]) {
  out[0] = (real)(c & 0xFFu) / 255.0f;
  out[1] = (real)((c >> 8) & 0xFFu) / 255.0f;
  out[2] = (real)((c >> 16) & 0xFFu) / 255.0f;
  out[3] = (real)((c >> 24) & 0xFFu) / 255.0f;
}

// ---- trace.rs ----
struct Hit {
  real t;
  uint32_t quad, texel;
  real s, tt;
};
static const Hit NO_HIT = {INFINITY, UINT32_MAX, 0u, 0.0f, 0.0f};
static const uint32_t MAX_STEPS = 65536u;
static const real COPLANAR = 1e-5f;

// trace.rs:37-58
static inline bool intersect_quad(const Quad &q, V3 o, V3 d, real tmin, real tmax, real &t_out,
                                  real &s_out, real &tt_out) {
  V3 eu = from4(q.eu), ev = from4(q.ev);
  V3 n = cross(eu, ev);
  real denom = dot(n, d);
  if (f_abs(denom) < 1e-12f) return false;
  V3 p0 = from4(q.p0);
  real t = dot(n, p0 - o) / denom;
  if (!(t > tmin && t < tmax)) return false;
  V3 p = o + d * t - p0;
  real nn = dot(n, n);
  real s = dot(cross(p, ev), n) / nn;
  real tt = dot(cross(eu, p), n) / nn;
  if (s < 0.0f || s > 1.0f || tt < 0.0f || tt > 1.0f) return false;
  t_out = t;
  s_out = s;
  tt_out = tt;
  return true;
}

// trace.rs:62-69
static inline uint32_t quad_texel(const View &s, const Quad &q, real ss, real t) {
  ss = f_min(f_max(ss, 0.0005f), 0.9995f);
  t = f_min(f_max(t, 0.0005f), 0.9995f);
  real u = q.uv[0] + ss * q.uv[2] + t * q.uvt[0];
  real v = q.uv[1] + ss * q.uv[3] + t * q.uvt[1];
  return texel(s, u, v);
}

// trace.rs:81-109
static inline void test_quad(const View &s, uint32_t qi, V3 o, V3 d, real tmin, Hit &best) {
  const Quad &q = quad_at(s, qi);
  LWO_COUNT(bytes, 80);
  LWO_COUNT(quads_tested, 1);
  real reach = (best.quad != UINT32_MAX && qi > best.quad) ? best.t + COPLANAR : best.t;
  real t, ss, tt;
  if (intersect_quad(q, o, d, tmin, reach, t, ss, tt)) {
    if ((q_flags(q) & FLAG_FRONT_ONLY) != 0 && dot(q_normal(q), d) >= 0.0f) return;
    uint32_t tx = quad_texel(s, q, ss, tt);
    uint32_t alpha = tx >> 24;
    uint32_t mode = q_flags(q) & FLAG_ALPHA_MASK;
    if (mode == ALPHA_TRANSLUCENT && alpha == 0u) LWO_COUNT(alpha_zero_rejects, 1);
    if ((mode == ALPHA_CUTOUT && alpha < 128u) || (mode == ALPHA_TRANSLUCENT && alpha == 0u)) return;
    best = Hit{t, qi, tx, ss, tt};
  }
}

// trace.rs:113-116
static inline bool is_see_through(const View &s, uint32_t qi) {
  const uint32_t *st = s.params->see_through;
  return qi >= st[0] && qi - st[0] < st[1];
}

// trace.rs:120-131
static inline bool see_through_hit(const View &s, V3 o, V3 d, real tmin, Hit &out) {
  const uint32_t *st = s.params->see_through;
  Hit best = NO_HIT;
  for (uint32_t i = 0; i < st[1]; i++) test_quad(s, st[0] + i, o, d, tmin, best);
  if (best.quad != UINT32_MAX) {
    out = best;
    return true;
  }
  return false;
}

// trace.rs:293-319
static inline real safe_inv(real x) {
  if (f_abs(x) < 1e-12f) return x < 0.0f ? -1e12f : 1e12f;
  return 1.0f / x;
}
static inline int32_t sign_i(real x) { return x > 0.0f ? 1 : (x < 0.0f ? -1 : 0); }
static inline real next_boundary(real o, real inv, int32_t cell, int32_t step) {
  if (step == 0) return INFINITY;
  real bound = (real)(step > 0 ? cell + 1 : cell);
  return (bound - o) * inv;
}

static inline int32_t clamp_i(int32_t v, int32_t lo, int32_t hi) {
  if (lo > hi) throw OracleError("clamp with lo > hi");
  return v < lo ? lo : (v > hi ? hi : v);
}

// trace.rs:137-290; cell_out receives the cell the traversal stopped in.
static bool trace(const View &s, V3 o, V3 d, real tmin, real tmax, Hit &out,
                  int32_t cell_out[3]) {
  const Params &p = *s.params;
  const uint32_t *dims = p.grid_dims;
  V3 size = grid_size(s);
  LWO_COUNT(traces, 1);

  V3 inv = v3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
  V3 ta = mul(ZERO3 - o, inv);
  V3 tb = mul(size - o, inv);
  real t_enter = f_max(f_max(f_min(ta.x, tb.x), f_min(ta.y, tb.y)), f_min(ta.z, tb.z));
  real t_exit = f_min(f_min(f_max(ta.x, tb.x), f_max(ta.y, tb.y)), f_max(ta.z, tb.z));
  real t = f_max(t_enter, 0.0f);
  real t_end = f_min(t_exit, tmax);
  int32_t cell[3] = {0, 0, 0};
  cell_out[0] = cell_out[1] = cell_out[2] = -1;
  if (t > t_end) return false;

  V3 start = o + d * (t + 1e-4f);
  int32_t gs[3] = {(int32_t)dims[0] * 16, (int32_t)dims[1] * 16, (int32_t)dims[2] * 16};
  cell[0] = clamp_i(as_i32(f_floor(start.x)), 0, gs[0] - 1);
  cell[1] = clamp_i(as_i32(f_floor(start.y)), 0, gs[1] - 1);
  cell[2] = clamp_i(as_i32(f_floor(start.z)), 0, gs[2] - 1);
  int32_t step[3] = {sign_i(d.x), sign_i(d.y), sign_i(d.z)};
  real next[3];
  for (int a = 0; a < 3; a++) next[a] = next_boundary(get(o, a), get(inv, a), cell[a], step[a]);

  uint32_t top_off = p.off[0];
  uint32_t ov_top_off = p.off[2];
  real best_t = t_end + 1e-3f;

  for (uint32_t it = 0; it < MAX_STEPS; it++) {
    uint32_t bx = (uint32_t)(cell[0] >> 4);
    uint32_t by = (uint32_t)(cell[1] >> 4);
    uint32_t bz = (uint32_t)(cell[2] >> 4);
    uint32_t b = (by * dims[2] + bz) * dims[0] + bx;
    uint32_t entry = grid_at(s, (size_t)(top_off + b));
    uint32_t ov = grid_at(s, (size_t)(ov_top_off + b));
    LWO_COUNT(bytes, 8);  // the top-level and overlay entries
    LWO_COUNT(dda_steps, 1);
    cell_out[0] = cell[0];
    cell_out[1] = cell[1];
    cell_out[2] = cell[2];

    if (entry == EMPTY && ov == EMPTY) {
#if LWO_COUNTING
      if (tl_octant_boxes) {
        const uint32_t octant = (step[0] >= 0 ? 1u : 0u) | (step[1] >= 0 ? 2u : 0u) | (step[2] >= 0 ? 4u : 0u);
        if ((((*tl_octant_boxes)[b] >> octant) & 1u) == 0u) return false;
      }
#endif
      // An empty brick is crossed in one move, to the face where the ray leaves it.
      int32_t bc[3] = {(int32_t)bx, (int32_t)by, (int32_t)bz};
      real t_out = INFINITY;
      int axis = 0;
      for (int a = 0; a < 3; a++) {
        if (step[a] != 0) {
          int32_t bound = step[a] > 0 ? (bc[a] + 1) * 16 : bc[a] * 16;
          real tt = ((real)bound - get(o, a)) * get(inv, a);
          if (tt < t_out) {
            t_out = tt;
            axis = a;
          }
        }
      }
      if (t_out > t_end) break;
      V3 pos = o + d * t_out;
      for (int a = 0; a < 3; a++) {
        if (a == axis) {
          cell[a] = step[a] > 0 ? (bc[a] + 1) * 16 : bc[a] * 16 - 1;
        } else {
          cell[a] = clamp_i(as_i32(f_floor(get(pos, a))), bc[a] * 16, bc[a] * 16 + 15);
        }
      }
      if (cell[axis] < 0 || cell[axis] >= gs[axis]) break;
      for (int a = 0; a < 3; a++) next[a] = next_boundary(get(o, a), get(inv, a), cell[a], step[a]);
      t = t_out;
      continue;
    }

    real t_cell_exit = f_min(next[0], f_min(next[1], next[2]));
    real t_lo = f_max(t - 1e-4f, tmin);
    real t_hi = t_cell_exit + 1e-4f;

    uint32_t model;
    if (entry == EMPTY) {
      model = 0;
    } else if ((entry & UNIFORM) != 0) {
      model = entry & ~UNIFORM;
    } else {
      uint32_t local =
          (uint32_t)((((cell[1] & 15) * 16 + (cell[2] & 15)) * 16) + (cell[0] & 15));
      uint32_t w = grid_at(s, (size_t)(p.off[1] + entry * BRICK_WORDS + (local >> 1)));
      LWO_COUNT(bytes, 4);
      model = (w >> ((local & 1u) * 16u)) & 0xFFFFu;
    }
    Hit cand = NO_HIT;
    cand.t = f_min(t_hi, best_t);
    if (model != 0) {
      const Model &m = at(s.models, s.models_len, model, "models");
      LWO_COUNT(bytes, 16);
      V3 lo = o - v3((real)cell[0], (real)cell[1], (real)cell[2]);
      for (uint32_t i = 0; i < m.quad_count; i++) test_quad(s, m.quad_offset + i, lo, d, t_lo, cand);
    }
    if (ov != EMPTY) {
      uint32_t local =
          (uint32_t)((((cell[1] & 15) * 16 + (cell[2] & 15)) * 16) + (cell[0] & 15));
      size_t base = (size_t)(p.off[3] + ov * OV_CELL_WORDS + local * 2u);
      uint32_t off = grid_at(s, base);
      uint32_t count = grid_at(s, base + 1);
      LWO_COUNT(bytes, 8);
      for (uint32_t i = 0; i < count; i++) {
        uint32_t qi = grid_at(s, (size_t)(p.off2[0] + off + i));
        LWO_COUNT(bytes, 4);
        test_quad(s, qi, o, d, t_lo, cand);
      }
    }
    if (cand.quad != UINT32_MAX) {
      out = cand;
      return true;
    }

    int axis;
    if (next[0] < next[1]) {
      axis = next[0] < next[2] ? 0 : 2;
    } else if (next[1] < next[2]) {
      axis = 1;
    } else {
      axis = 2;
    }
    t = next[axis];
    if (t > t_end) break;
    cell[axis] += step[axis];
    if (cell[axis] < 0 || cell[axis] >= gs[axis]) break;
    next[axis] = next_boundary(get(o, axis), get(inv, axis), cell[axis], step[axis]);
  }
  return false;
}

static inline bool trace(const View &s, V3 o, V3 d, real tmin, real tmax, Hit &out) {
  int32_t cell[3];
  return trace(s, o, d, tmin, tmax, out, cell);
}

// ---- integrator.rs ----
static const uint32_t MAX_EVENTS = 48u;

// integrator.rs:26-34
static inline V3 linear_rgb(uint32_t c) {
  real u[4];
  unpack_rgba(c, u);
  return V3{gamma_to_linear(u[0]), gamma_to_linear(u[1]), gamma_to_linear(u[2])};
}

// integrator.rs:44-74
static V3 preetham_radiance(const View &s, V3 d) {
  const Params &p = *s.params;
  V3 sun = from4(p.sky_sun);
  real cos_theta = f_max(d.y + p.sky_zenith[3], 0.0f);
  real cos_gamma = f_clamp(dot(d, sun), -1.0f, 1.0f);
  real gamma = f_acos(cos_gamma);
  real cos2_gamma = cos_gamma * cos_gamma;
  auto perez = [&](int k) -> real {
    real a = cos_theta > 0.0f ? p.perez_a[k] * f_exp(p.perez_b[k] / cos_theta) : 0.0f;
    return (1.0f + a) *
           (1.0f + p.perez_c[k] * f_exp(p.perez_d[k] * gamma) + p.perez_e[k] * cos2_gamma);
  };
  real x = p.sky_zenith[0] * perez(0);
  real y = p.sky_zenith[1] * perez(1);
  real lum = p.sky_zenith[2] * perez(2);
  if (y <= 5e-8f) return ZERO3;
  real f = lum / y;
  real cx = x * f, cy = lum, cz = (1.0f - x - y) * f;
  return V3{2.3706743f * cx - 0.9000405f * cy - 0.4706338f * cz,
            -0.513885f * cx + 1.4253036f * cy + 0.0885814f * cz,
            0.0052982f * cx - 0.0146949f * cy + 1.0093968f * cz} *
         0.045f;
}

// integrator.rs:79-104
static inline void rgb_to_hsl(V3 c, real out[3]) {
  real r = f_clamp(c.x, 0.0f, 1.0f), g = f_clamp(c.y, 0.0f, 1.0f), b = f_clamp(c.z, 0.0f, 1.0f);
  real cmax = f_max(f_max(r, g), b);
  real cmin = f_min(f_min(r, g), b);
  real delta = cmax - cmin;
  real lightness = (cmax + cmin) / 2.0f;
  real hue;
  if (delta == 0.0f) {
    hue = 0.0f;
  } else if (cmax == r) {
    hue = f_rem((g - b) / delta, 6.0f) / 6.0f;
  } else if (cmax == g) {
    hue = (((b - r) / delta) + 2.0f) / 6.0f;
  } else {
    hue = (((r - g) / delta) + 4.0f) / 6.0f;
  }
  real saturation = delta < 5e-8f ? 0.0f : delta / (1.0f - f_abs(2.0f * lightness - 1.0f));
  out[0] = hue;
  out[1] = saturation;
  out[2] = lightness;
}

// integrator.rs:107-127
static inline V3 rgb_from_hsl(real hue, real saturation, real lightness) {
  real c = f_min(1.0f, (1.0f - f_abs(2.0f * lightness - 1.0f)) * saturation);
  real h = hue * 6.0f;
  real x = c * (1.0f - f_abs(f_rem(h, 2.0f) - 1.0f));
  V3 rgb;
  if (h < 1.0f) {
    rgb = v3(c, x, 0.0f);
  } else if (h < 2.0f) {
    rgb = v3(x, c, 0.0f);
  } else if (h < 3.0f) {
    rgb = v3(0.0f, c, x);
  } else if (h < 4.0f) {
    rgb = v3(0.0f, x, c);
  } else if (h < 5.0f) {
    rgb = v3(x, 0.0f, c);
  } else {
    rgb = v3(c, 0.0f, x);
  }
  real m = f_max(0.0f, lightness - 0.5f * c);
  return rgb + v3(m, m, m);
}

// integrator.rs:133-148
static void build_sky_cache(const View &s, uint32_t n, float *out, size_t out_len) {
  uint32_t m = n + 1;
  for (uint32_t i = 0; i < m; i++) {
    real sin_t, cos_t;
    sin_cos_turns((real)i / (real)n, sin_t, cos_t);
    for (uint32_t j = 0; j < m; j++) {
      real sin_p, cos_p;
      sin_cos_turns((real)j / (real)(2u * n) - 0.25f, sin_p, cos_p);
      real r = f_abs(cos_p);
      real hsl[3];
      rgb_to_hsl(preetham_radiance(s, v3(cos_t * r, sin_p, sin_t * r)), hsl);
      size_t a = (size_t)((i * m + j) * 3u);
      if (a + 2 >= out_len) throw OracleError("sky cache output too small");
      out[a] = hsl[0];
      out[a + 1] = hsl[1];
      out[a + 2] = hsl[2];
    }
  }
}

static inline real sky_at(const View &s, size_t i) { return at(s.sky, s.sky_len, i, "sky"); }

// integrator.rs:158-188
static V3 sky_radiance(const View &s, V3 d) {
  const Params &p = *s.params;
  real n = p.misc_f[2];
  if (n < 1.0f) return ZERO3;
  LWO_COUNT(bytes, 48);
  LWO_COUNT(sky_lookups, 1);
  real theta = f_atan2(d.z, d.x) / (2.0f * PI_F);
  theta = f_rem(f_rem(theta, 1.0f) + 1.0f, 1.0f);
  real phi = 1.0f - f_acos(f_clamp(d.y, -1.0f, 1.0f)) / PI_F;
  real x = theta * n, y = phi * n;
  real fx = f_floor(f_clamp(x, 0.0f, n - 1.0f));
  real fy = f_floor(f_clamp(y, 0.0f, n - 1.0f));
  uint32_t m = as_u32(n) + 1u;
  auto idx = [&](real i, real j) -> size_t { return (size_t)((as_u32(i) * m + as_u32(j)) * 3u); };
  size_t a00 = idx(fx, fy), a10 = idx(fx + 1.0f, fy), a01 = idx(fx, fy + 1.0f),
         a11 = idx(fx + 1.0f, fy + 1.0f);
  real hsl[3];
  for (int k = 0; k < 3; k++) {
    real y0 = sky_at(s, a00 + k) + (x - fx) * (sky_at(s, a10 + k) - sky_at(s, a00 + k));
    real y1 = sky_at(s, a01 + k) + (x - fx) * (sky_at(s, a11 + k) - sky_at(s, a01 + k));
    hsl[k] = y0 + (y - fy) * (y1 - y0);
  }
  return rgb_from_hsl(hsl[0], hsl[1], hsl[2]) * p.sky_sun[3];
}

// integrator.rs:196-217
static V3 map_lookup(const View &s, uint32_t base, real w, real h, V3 d) {
  LWO_COUNT(bytes, 48);
  LWO_COUNT(sky_lookups, 1);
  real u = f_atan2(d.z, d.x) / (2.0f * PI_F);
  u = f_rem(f_rem(u, 1.0f) + 1.0f, 1.0f);
  real v = 1.0f - f_acos(f_clamp(d.y, -1.0f, 1.0f)) / PI_F;
  real x = u * w - 0.5f;
  real y = f_clamp(v * h - 0.5f, 0.0f, h - 1.0f);
  real fx = f_floor(x);
  real fy = f_floor(y);
  real tx = x - fx, ty = y - fy;
  real i0 = fx < 0.0f ? w - 1.0f : fx;
  real i1 = fx + 1.0f >= w ? 0.0f : fx + 1.0f;
  real j1 = f_min(fy + 1.0f, h - 1.0f);
  auto texel_at = [&](real i, real j) -> V3 {
    size_t k = (size_t)(base + (as_u32(j) * as_u32(w) + as_u32(i)) * 3u);
    return v3(sky_at(s, k), sky_at(s, k + 1), sky_at(s, k + 2));
  };
  V3 top = texel_at(i0, fy) * (1.0f - tx) + texel_at(i1, fy) * tx;
  V3 bottom = texel_at(i0, j1) * (1.0f - tx) + texel_at(i1, j1) * tx;
  return top * (1.0f - ty) + bottom * ty;
}

// integrator.rs:225-231
static V3 sky_lens(const View &s, V3 d) {
  const float *m = s.params->sky_map;
  if (m[0] < 1.0f) return sky_radiance(s, d);
  return map_lookup(s, 0, m[0], m[1], d);
}

// integrator.rs:237-243
static V3 sky_light(const View &s, V3 d) {
  const float *m = s.params->sky_map;
  if (m[0] < 1.0f) return sky_radiance(s, d);
  return map_lookup(s, as_u32(m[0] * m[1]) * 3u, m[2], m[3], d);
}

// integrator.rs:252-277
static V3 sun_texture(const View &s, V3 d) {
  const Params &p = *s.params;
  const float *sprite = p.sun_sprite;
  if (sprite[2] <= 0.0f || dot(d, from4(p.sun_dir)) < 0.5f) return ZERO3;
  real width = p.sun_sv[3];
  real width2 = width * 2.0f;
  real a = 0.5f * PI_F - f_acos(dot(d, from4(p.sun_su))) + width;
  real b = 0.5f * PI_F - f_acos(dot(d, from4(p.sun_sv))) + width;
  if (a < 0.0f || a >= width2 || b < 0.0f || b >= width2) return ZERO3;
  real x = f_clamp(f_floor(a / width2 * sprite[2] - 5e-8f), 0.0f, sprite[2] - 1.0f);
  real y = f_clamp(f_floor((1.0f - b / width2) * sprite[3] - 5e-8f), 0.0f, sprite[3] - 1.0f);
  return linear_rgb(texel(s, sprite[0] + x + 0.5f, sprite[1] + y + 0.5f)) * p.sun_su[3];
}

// integrator.rs:279-287
static inline V3 cosine_hemisphere(V3 n, Rng &rng) {
  real r1 = rng.next();
  real r2 = rng.next();
  real r = f_sqrt(r1);
  real sin_phi, cos_phi;
  sin_cos_turns(r2, sin_phi, cos_phi);
  V3 t, b;
  onb(n, t, b);
  return normalize(t * (r * cos_phi) + b * (r * sin_phi) + n * f_sqrt(f_max(1.0f - r1, 0.0f)));
}

// integrator.rs:289-302
static inline V3 sample_sun(const View &s, Rng &rng) {
  const Params &p = *s.params;
  V3 sun = from4(p.sun_dir);
  real cos_max = p.sun_dir[3];
  real r1 = rng.next();
  real r2 = rng.next();
  real cos_t = 1.0f - r1 * (1.0f - cos_max);
  real sin_t = f_sqrt(f_max(1.0f - cos_t * cos_t, 0.0f));
  real sin_phi, cos_phi;
  sin_cos_turns(r2, sin_phi, cos_phi);
  V3 t, b;
  onb(sun, t, b);
  return normalize(t * (sin_t * cos_phi) + b * (sin_t * sin_phi) + sun * cos_t);
}

// integrator.rs:307-315
static inline V3 shading_normal(const Quad &q, V3 d, bool &front) {
  V3 n = normalize(q_normal(q));
  if (dot(n, d) < 0.0f) {
    front = true;
    return n;
  }
  front = false;
  return -n;
}

// integrator.rs:318-346
static V3 shadow(const View &s, V3 o, V3 d, real tmax) {
  real eps = s.params->misc_f[1];
  V3 tr = ONE3;
  for (int i = 0; i < 8; i++) {
    Hit h;
    if (!trace(s, o, d, eps, tmax, h)) return tr;
    const Quad &q = quad_at(s, h.quad);
    LWO_COUNT(bytes, 80);
    uint32_t flags = q_flags(q);
    if ((flags & FLAG_WATER) != 0) {
      if (dot(q_normal(q), d) > 0.0f) {
        LWO_COUNT(shadow_water_exits, 1);
        tr = mul(tr, transmittance(from4(s.params->water_color), h.t));
      }
    } else if ((flags & FLAG_ALPHA_MASK) == ALPHA_TRANSLUCENT) {
      LWO_COUNT(shadow_translucent, 1);
      real a = (real)(h.texel >> 24) / 255.0f;
      V3 col = mul(linear_rgb(h.texel), linear_rgb(q_tint(q)));
      tr = mul(tr, lerp(ONE3, col, a));
    } else {
      return ZERO3;
    }
    o = o + d * (h.t + eps);
    tmax -= h.t + eps;
  }
  return tr;
}

// integrator.rs:349-353
static inline uint32_t pick(uint32_t n, real u) {
  uint32_t i = as_u32(u * (real)n);
  return i < n ? i : n - 1;
}

static inline uint32_t eg_at(const View &s, size_t i) {
  return at(s.emitter_grid, s.emitter_grid_len, i, "emitter_grid");
}

// integrator.rs:357-373
static inline void emitter_cell(const View &s, V3 p, uint32_t &start, uint32_t &count) {
  real cell = (real)eg_at(s, 0);
  int32_t size[3] = {(int32_t)eg_at(s, 1), (int32_t)eg_at(s, 2), (int32_t)eg_at(s, 3)};
  int32_t c[3] = {
      (int32_t)((uint32_t)as_i32(f_floor(p.x / cell)) - eg_at(s, 4)),
      (int32_t)((uint32_t)as_i32(f_floor(p.y / cell)) - eg_at(s, 5)),
      (int32_t)((uint32_t)as_i32(f_floor(p.z / cell)) - eg_at(s, 6)),
  };
  LWO_COUNT(bytes, 28);
  if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= size[0] || c[1] >= size[1] || c[2] >= size[2]) {
    start = 0;
    count = 0;
    return;
  }
  uint32_t i = ((uint32_t)c[1] * (uint32_t)size[2] + (uint32_t)c[2]) * (uint32_t)size[0] +
               (uint32_t)c[0];
  size_t a = (size_t)(EMITTER_GRID_HEADER + i * 2u);
  start = eg_at(s, a);
  count = eg_at(s, a + 1);
  LWO_COUNT(bytes, 8);
}

// integrator.rs:380-414
static V3 sample_emitter_face(const View &s, V3 origin, V3 n, const Emitter &e, uint32_t face,
                              real scaler, Rng &rng) {
  const Params &p = *s.params;
  const Quad &q = quad_at(s, (size_t)(f_bits(e.pos[3]) + face));
  LWO_COUNT(bytes, 80);
  real ss = rng.next();
  real t = rng.next();
  V3 at_p = from4(e.pos) + from4(q.p0) + from4(q.eu) * ss + from4(q.ev) * t;
  V3 to = at_p - origin;
  if (dot(to, n) <= 0.0f) return ZERO3;
  real dist = length(to);
  V3 l = to / dist;
  real eps = p.misc_f[1];
  Hit h;
  if (!trace(s, origin, l, eps, dist + 2.0f * eps, h)) return ZERO3;
  if (f_abs(h.t - dist) >= 2.0f * eps) return ZERO3;
  V3 nq = q_normal(q);
  real area = length(nq);
  real cos_l = f_abs(dot(nq, l)) / area;
  const Quad &hq = quad_at(s, h.quad);
  V3 c = mul(linear_rgb(h.texel), linear_rgb(q_tint(hq)));
  return c * (cos_l / f_max(dist * dist, 1.0f) * area * q_emittance(q) * p.misc_f[0] * scaler);
}

// integrator.rs:418-433
static V3 sample_emitter(const View &s, V3 origin, V3 n, const Emitter &e, Rng &rng) {
  uint32_t faces = e.faces[0];
  if (faces == 0) return ZERO3;
  if (s.params->emitter[0] == EMITTER_SAMPLING_ONE) {
    uint32_t f = pick(faces, rng.next());
    return sample_emitter_face(s, origin, n, e, f, 1.0f, rng);
  }
  V3 sum = ZERO3;
  for (uint32_t f = 0; f < faces; f++)
    sum = sum + sample_emitter_face(s, origin, n, e, f, 1.0f / (real)faces, rng);
  return sum;
}

// integrator.rs:441-469
static V3 sample_emitters(const View &s, V3 pos, V3 origin, V3 n, Rng &rng) {
  uint32_t start, count;
  emitter_cell(s, pos, start, count);
  if (count == 0) return ZERO3;
  size_t list = (size_t)EMITTER_GRID_HEADER +
                (size_t)eg_at(s, 1) * (size_t)eg_at(s, 2) * (size_t)eg_at(s, 3) * 2 + (size_t)start;
  uint32_t strategy = s.params->emitter[0];
  if (strategy == EMITTER_SAMPLING_ONE || strategy == EMITTER_SAMPLING_ONE_BLOCK) {
    uint32_t i = eg_at(s, list + pick(count, rng.next()));
    LWO_COUNT(bytes, 4 + 32);
    Emitter e = at(s.emitters, s.emitters_len, i, "emitters");
    return sample_emitter(s, origin, n, e, rng) * PI_F;
  }
  if (strategy == EMITTER_SAMPLING_ALL) {
    V3 sum = ZERO3;
    for (uint32_t k = 0; k < count; k++) {
      uint32_t i = eg_at(s, list + k);
      LWO_COUNT(bytes, 4 + 32);
      Emitter e = at(s.emitters, s.emitters_len, i, "emitters");
      sum = sum + sample_emitter(s, origin, n, e, rng);
    }
    return sum * (PI_F / (real)count);
  }
  return ZERO3;
}

// integrator.rs:477-505
static inline void water_interface(const Params &p, V3 pos, V3 d, V3 n, bool front, bool in_water,
                                   real eps, Rng &rng, V3 &d_out, V3 &o_out, bool &w_out) {
  real ior = p.water_color[3];
  real eta = front ? 1.0f / ior : ior;
  real cos_i = -dot(n, d);
  real sin2_t = eta * eta * f_max(1.0f - cos_i * cos_i, 0.0f);
  real r0 = (1.0f - ior) / (1.0f + ior);
  r0 = r0 * r0;
  real m = 1.0f - cos_i;
  real fresnel = r0 + (1.0f - r0) * m * m * m * m * m;
  if (sin2_t >= 1.0f || rng.next() < fresnel) {
    d_out = normalize(d + n * (2.0f * cos_i));
    o_out = pos + n * eps;
    w_out = in_water;
  } else {
    real cos_t = f_sqrt(1.0f - sin2_t);
    d_out = normalize(d * eta + n * (eta * cos_i - cos_t));
    o_out = pos - n * eps;
    w_out = front;
  }
}

// integrator.rs:510-526
static inline bool water_plane(const View &s, V3 o, V3 d, real &t_out) {
  const Params &p = *s.params;
  if (p.water_plane[1] == 0.0f || d.y == 0.0f) return false;
  real t = (p.water_plane[0] - o.y) / d.y;
  if (!(t > 0.0f && t < INFINITY)) return false;
  V3 a = o + d * t;
  V3 size = grid_size(s);
  if (a.x >= 0.0f && a.z >= 0.0f && a.x < size.x && a.z < size.z) return false;
  t_out = t;
  return true;
}

// integrator.rs:530-534
static inline bool under_water_plane(const View &s, V3 o, V3 d) {
  const Params &p = *s.params;
  return p.water_plane[1] != 0.0f && o.y < p.water_plane[0] && d.y <= 0.0f;
}

// integrator.rs:539-709
static V3 render_sample(const View &s, uint32_t px, uint32_t py, uint32_t sample) {
  const Params &p = *s.params;
  uint32_t width = p.frame[0];
  uint32_t height = p.frame[1];
  uint32_t max_bounces = p.frame[3];
  real eps = p.misc_f[1];
  Rng rng = Rng::make(py * width + px, sample, p.misc_u[0]);
  LWO_COUNT(samples, 1);

  real tan_half = p.cam_pos[3];
  real aspect = p.cam_fwd[3];
  real jx = rng.next();
  real jy = rng.next();
  real sx = (2.0f * ((real)px + jx) / (real)width - 1.0f) * aspect * tan_half;
  real sy = (1.0f - 2.0f * ((real)py + jy) / (real)height) * tan_half;
  V3 d = normalize(from4(p.cam_fwd) + from4(p.cam_right) * sx + from4(p.cam_up) * sy);
  V3 o = from4(p.cam_pos);

  V3 radiance = ZERO3;
  V3 beta = ONE3;
  uint32_t bounces = 0;
  bool specular = true;
  bool in_water = false;
  uint32_t strategy = p.misc_u[3] > 0 ? p.emitter[0] : EMITTER_SAMPLING_NONE;
  real tmin = 0.0f;
  V3 sigma = from4(p.water_color);

  for (uint32_t ev = 0; ev < MAX_EVENTS; ev++) {
    Hit hit;
    bool has_hit = trace(s, o, d, tmin, INFINITY, hit);
    bool forced = false;
    if (bounces == 0 && p.see_through[1] > 0 && has_hit && !is_see_through(s, hit.quad)) {
      Hit st;
      if (see_through_hit(s, o, d, hit.t, st)) {
        const Quad &sq = quad_at(s, st.quad);
        if ((q_flags(sq) & FLAG_ALPHA_MASK) == ALPHA_TRANSLUCENT) {
          if (rng.next() < (real)(st.texel >> 24) / 255.0f) {
            hit = st;
            forced = true;
            LWO_COUNT(see_through_drawn, 1);
            LWO_COUNT(see_through_forced, 1);
          }
        } else {
          hit = st;
          LWO_COUNT(see_through_drawn, 1);
        }
      }
    }
    real tp;
    if (water_plane(s, o, d, tp) && (!has_hit || tp < hit.t)) {
      if (in_water) beta = mul(beta, transmittance(sigma, tp));
      bool front = d.y < 0.0f;
      V3 n = v3(0.0f, front ? 1.0f : -1.0f, 0.0f);
      water_interface(p, o + d * tp, d, n, front, in_water, eps, rng, d, o, in_water);
      specular = true;
      tmin = 0.0f;
      continue;
    }
    if (!has_hit) {
      if (!in_water && !under_water_plane(s, o, d)) {
        if (bounces == 0) {
          radiance = radiance + mul(beta, sky_lens(s, d));
        } else {
          radiance = radiance + mul(beta, sky_light(s, d));
        }
        if (specular && bounces == 0) radiance = radiance + mul(beta, sun_texture(s, d));
      }
      break;
    }
    const Hit &h = hit;
    if (in_water) beta = mul(beta, transmittance(sigma, h.t));
    const Quad &q = quad_at(s, h.quad);
    LWO_COUNT(bytes, 80);
    uint32_t flags = q_flags(q);
    V3 pos = o + d * h.t;
    real alpha = (real)(h.texel >> 24) / 255.0f;
    V3 albedo = mul(linear_rgb(h.texel), linear_rgb(q_tint(q)));
    bool front;
    V3 n = shading_normal(q, d, front);

    if ((flags & FLAG_WATER) != 0) {
      LWO_COUNT(water_interfaces, 1);
      water_interface(p, pos, d, n, front, in_water, eps, rng, d, o, in_water);
      specular = true;
      tmin = 0.0f;
      continue;
    }

    if (!forced && (flags & FLAG_ALPHA_MASK) == ALPHA_TRANSLUCENT && rng.next() >= alpha) {
      LWO_COUNT(translucent_passes, 1);
      beta = mul(beta, lerp(ONE3, albedo, alpha));
      o = pos + d * eps;
      tmin = 0.0f;
      continue;
    }

    if (bounces >= max_bounces) break;
    real emittance = q_emittance(q) * p.misc_f[0];
    bool emits = emittance > 0.0f &&
                 (p.emitter[1] == 0 || strategy == EMITTER_SAMPLING_NONE || bounces == 0);
    if (emits) radiance = radiance + mul(beta, mul(albedo, albedo)) * emittance;
    bounces += 1;

    V3 origin = pos + n * eps;
    V3 sun = from4(p.sun_dir);
    real sin_r = f_sqrt(f_max(1.0f - p.sun_dir[3] * p.sun_dir[3], 0.0f));
    bool lit = p.sun_color[0] > 0.0f || p.sun_color[1] > 0.0f || p.sun_color[2] > 0.0f;
    if (lit && dot(n, sun) > -sin_r && p.sun_dir[1] > -sin_r) {
      V3 l = sample_sun(s, rng);
      real cos_l = dot(n, l);
      if (cos_l > 0.0f && l.y >= 0.0f) {
        V3 tr = shadow(s, origin, l, INFINITY);
        radiance = radiance + mul(mul(mul(beta, albedo), from4(p.sun_color)), tr) * cos_l;
      }
    }
    if (!emits && strategy != EMITTER_SAMPLING_NONE) {
      radiance = radiance + mul(mul(beta, albedo), sample_emitters(s, pos, origin, n, rng));
    }

    beta = mul(beta, albedo);
    if (bounces >= 3) {
      real q_rr = f_clamp(max_elem(beta), 0.05f, 0.95f);
      if (rng.next() > q_rr) break;
      beta = beta / q_rr;
    }
    d = cosine_hemisphere(n, rng);
    o = origin;
    tmin = 0.0f;
    specular = false;
  }
  return radiance;
}

static inline bool finite3(V3 c) {
  return std::isfinite(c.x) && std::isfinite(c.y) && std::isfinite(c.z);
}

#if LWO_COUNTING
// tl_octant_boxes for a scene: each brick's bits by brute force over the bricks that are
// not EMPTY, each of which sets the octant bits of every brick whose box it lies in.
static std::vector<uint8_t> octant_boxes(const View &s) {
  const Params &p = *s.params;
  const int64_t dx = p.grid_dims[0], dy = p.grid_dims[1], dz = p.grid_dims[2];
  std::vector<uint8_t> boxes((size_t)(dx * dy * dz), 0);
  for (int64_t y = 0; y < dy; y++)
    for (int64_t z = 0; z < dz; z++)
      for (int64_t x = 0; x < dx; x++) {
        const size_t b = (size_t)((y * dz + z) * dx + x);
        if (grid_at(s, (size_t)p.off[0] + b) == EMPTY && grid_at(s, (size_t)p.off[2] + b) == EMPTY) continue;
        // brick (x, y, z) lies in the +x box of bricks with bx <= x, the -x box of bx >= x, ...
        for (int64_t by = 0; by < dy; by++)
          for (int64_t bz = 0; bz < dz; bz++)
            for (int64_t bx = 0; bx < dx; bx++) {
              uint8_t &v = boxes[(size_t)((by * dz + bz) * dx + bx)];
              for (uint32_t k = 0; k < 8; k++) {
                const bool in_x = (k & 1u) ? bx <= x : bx >= x;
                const bool in_y = (k & 2u) ? by <= y : by >= y;
                const bool in_z = (k & 4u) ? bz <= z : bz >= z;
                if (in_x && in_y && in_z) v |= (uint8_t)(1u << k);
              }
            }
      }
  return boxes;
}
#endif

// CpuBackend::render_hq (crates/lumenwright/src/render/cpu/mod.rs:39-87): rows in
// parallel (one row per task, taken dynamically as rayon's work stealing does), f64
// per-pixel accumulation of the finite samples, divided by spp.
// With counters, device_walk chooses the walk counted: the device's, which stops once every
// brick left toward the ray's octant is EMPTY, or (false) the reference's whole walk.
static void render_hq(const View &view_in, uint32_t spp_in, unsigned threads, real *out,
                      Counters *counters, bool device_walk = true) {
  Params params = *view_in.params;
  params.frame[2] = 0;
  View view = view_in;
  view.params = &params;
  const uint32_t w = params.frame[0], h = params.frame[1];
  const uint32_t spp = spp_in > 0 ? spp_in : 1;
  std::atomic<uint32_t> next_row{0};
  std::vector<std::string> errors(threads);
  std::vector<Counters> per_thread(threads);
#if LWO_COUNTING
  const std::vector<uint8_t> boxes = octant_boxes(view);
#endif
  auto worker = [&](unsigned tid) {
#if LWO_COUNTING
    per_thread[tid] = Counters{};
    tl_counters = counters ? &per_thread[tid] : nullptr;
    tl_octant_boxes = counters && device_walk ? &boxes : nullptr;
#endif
    try {
      for (;;) {
        uint32_t y = next_row.fetch_add(1);
        if (y >= h) break;
        for (uint32_t x = 0; x < w; x++) {
          double acc[3] = {0.0, 0.0, 0.0};
          for (uint32_t sample = 0; sample < spp; sample++) {
            V3 c = render_sample(view, x, y, sample);
            if (finite3(c)) {
              acc[0] += (double)c.x;
              acc[1] += (double)c.y;
              acc[2] += (double)c.z;
            }
          }
          real *px = out + ((size_t)y * w + x) * 3;
          px[0] = (real)(acc[0] / (double)spp);
          px[1] = (real)(acc[1] / (double)spp);
          px[2] = (real)(acc[2] / (double)spp);
        }
      }
    } catch (const std::exception &e) {
      errors[tid] = e.what();
      next_row.store(h);
    }
#if LWO_COUNTING
    tl_counters = nullptr;
    tl_octant_boxes = nullptr;
#endif
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < threads; t++) pool.emplace_back(worker, t);
  worker(0);
  for (auto &t : pool) t.join();
  for (auto &e : errors)
    if (!e.empty()) throw OracleError(e);
#if LWO_COUNTING
  if (counters) {
    for (auto &c : per_thread) {
      counters->bytes += c.bytes;
      counters->dda_steps += c.dda_steps;
      counters->quads_tested += c.quads_tested;
      counters->traces += c.traces;
      counters->samples += c.samples;
      counters->sky_lookups += c.sky_lookups;
      counters->see_through_drawn += c.see_through_drawn;
      counters->see_through_forced += c.see_through_forced;
      counters->translucent_passes += c.translucent_passes;
      counters->alpha_zero_rejects += c.alpha_zero_rejects;
      counters->shadow_water_exits += c.shadow_water_exits;
      counters->shadow_translucent += c.shadow_translucent;
      counters->water_interfaces += c.water_interfaces;
    }
  }
#else
  (void)counters;
  (void)device_walk;
#endif
}

}  // namespace LWO_NS
