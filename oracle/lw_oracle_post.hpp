// ORACLE -- test infrastructure only. A restatement of the reference's post-processing,
// crates/lumenwright/src/render/post.rs (tone curves, tonemap, tonemap_mix, downsample), and
// of its PFM writer, crates/lumenwright-cli/src/output.rs:19-30, with the same f32 and f64
// operations in the same order and libm's powf (Rust std's f32::powf on Linux).
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace lwo_post {

// post.rs:3-21
enum Tonemap : uint32_t { GAMMA_OP = 0, ACES = 1, FILMIC = 2, FILMIC_MID_GREY = 3 };

static inline float rmax(float a, float b) {  // Rust f32::max: the non-NaN operand
  if (std::isnan(a)) return b;
  if (std::isnan(b)) return a;
  return a > b ? a : b;
}
static inline float rclamp(float x, float lo, float hi) {  // Rust f32::clamp (NaN stays)
  if (x < lo) return lo;
  if (x > hi) return hi;
  return x;
}

// post.rs:23-26
static inline float aces(float x) {
  const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
  return rclamp((x * (a * x + b)) / (x * (c * x + d) + e), 0.0f, 1.0f);
}

// post.rs:37-131
struct HableCurve {
  float a, b, c, d, e, f, w;
  float white_scale() const {
    float num = w * (a * w + c * b) + d * e;
    float den = w * (a * w + b) + d * f;
    return 1.0f / (num / den - e / f);
  }
  double h(double y) const {
    double aa = (double)a, bb = (double)b;
    double cb = (double)(c * b), de = (double)(d * e), df = (double)(d * f), ef = (double)(e / f);
    return (y * (aa * y + cb) + de) / (y * (aa * y + bb) + df) - ef;
  }
  double before_gamma(double x) const { return h(2.0 * x) * (double)white_scale(); }
  HableCurve brightened(double k) const {
    HableCurve r = *this;
    r.a = (float)((double)a * k * k);
    r.b = (float)((double)b * k);
    r.w = (float)((double)w / k);
    return r;
  }
  double mid_grey_gain(double m) const {
    double aa = (double)a, bb = (double)b;
    double cb = (double)(c * b), de = (double)(d * e), df = (double)(d * f), ef = (double)(e / f);
    double q = m / (double)white_scale() + ef;
    double qa = aa * (1.0 - q), qb = cb - q * bb, qc = de - q * df;
    double y = (-qb + std::sqrt(qb * qb - 4.0 * qa * qc)) / (2.0 * qa);
    return y / (2.0 * m);
  }
};

static const HableCurve HABLE_FILMIC_WORLDS = {0.15f, 0.5f, 0.1f, 0.2f, 0.02f, 0.3f, 11.2f};
static const double MID_GREY = 0.18;

// post.rs:136-138
static inline HableCurve game_tone_curve() {
  return HABLE_FILMIC_WORLDS.brightened(HABLE_FILMIC_WORLDS.mid_grey_gain(MID_GREY));
}

// post.rs:141-144 (GAMMA = 2.2, math.rs:171-176)
static inline float encode(float v) { return std::pow(rclamp(v, 0.0f, 1.0f), 1.0f / 2.2f); }

static inline HableCurve curve_of(uint32_t op) {
  return op == FILMIC_MID_GREY ? game_tone_curve() : HABLE_FILMIC_WORLDS;
}

// The operator's display value before the gamma (post.rs:155-165).
static inline float shown_linear(uint32_t op, const HableCurve &curve, double white, float v,
                                 float exposure) {
  v = rmax(v * exposure, 0.0f);
  switch (op) {
    case GAMMA_OP: return v;
    case ACES: return aces(v);
    default: return (float)(curve.h(2.0 * (double)v) * white);
  }
}

// Rust `as u32` of a float: saturating, NaN -> 0.
static inline uint32_t as_u32_sat(float x) {
  if (!(x > 0.0f)) return 0;
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

// post.rs:148-170
static void tonemap(const float *hdr, size_t n, float exposure, uint32_t op, uint32_t *out) {
  const HableCurve curve = curve_of(op);
  const double white = (double)curve.white_scale();
  for (size_t i = 0; i < n; i++) {
    uint32_t px = 0xFF000000u;
    for (int c = 0; c < 3; c++) {
      float v = shown_linear(op, curve, white, hdr[3 * i + c], exposure);
      px |= as_u32_sat(encode(v) * 255.0f + 0.5f) << (8 * c);
    }
    out[i] = px;
  }
}

// post.rs:176-215
static void tonemap_mix(const float *hdr, size_t n, float exposure, uint32_t a, uint32_t b,
                        float weight, uint32_t *out) {
  if (weight >= 1.0f) return tonemap(hdr, n, exposure, a, out);
  if (weight <= 0.0f) return tonemap(hdr, n, exposure, b, out);
  const HableCurve ca = curve_of(a), cb = curve_of(b);
  const double wa = (double)ca.white_scale(), wb = (double)cb.white_scale();
  for (size_t i = 0; i < n; i++) {
    uint32_t px = 0xFF000000u;
    for (int c = 0; c < 3; c++) {
      float v = hdr[3 * i + c];
      float d = weight * encode(shown_linear(a, ca, wa, v, exposure)) +
                (1.0f - weight) * encode(shown_linear(b, cb, wb, v, exposure));
      px |= as_u32_sat(d * 255.0f + 0.5f) << (8 * c);
    }
    out[i] = px;
  }
}

// post.rs:218-247
static std::vector<uint32_t> downsample(const uint32_t *px, uint32_t w, uint32_t h, uint32_t k) {
  if (k <= 1) return std::vector<uint32_t>(px, px + (size_t)w * h);
  uint32_t ow = w / k, oh = h / k;
  std::vector<uint32_t> out;
  out.reserve((size_t)ow * oh);
  for (uint32_t y = 0; y < oh; y++)
    for (uint32_t x = 0; x < ow; x++) {
      uint32_t acc[4] = {0, 0, 0, 0};
      for (uint32_t dy = 0; dy < k; dy++)
        for (uint32_t dx = 0; dx < k; dx++) {
          uint32_t c = px[(size_t)(y * k + dy) * w + x * k + dx];
          for (int i = 0; i < 4; i++) acc[i] += (c >> (8 * i)) & 0xFFu;
        }
      uint32_t n = k * k, v = 0;
      for (int i = 0; i < 4; i++) v += ((acc[i] + n / 2) / n) << (8 * i);
      out.push_back(v);
    }
  return out;
}

// output.rs:19-30: "PF\n{w} {h}\n-1.0\n" then the rows bottom to top, f32 little-endian.
static std::string pfm(const float *hdr, uint32_t w, uint32_t h) {
  std::string out = "PF\n" + std::to_string(w) + " " + std::to_string(h) + "\n-1.0\n";
  for (uint32_t yy = 0; yy < h; yy++) {
    uint32_t y = h - 1 - yy;
    out.append(reinterpret_cast<const char *>(hdr + (size_t)y * w * 3), (size_t)w * 3 * 4);
  }
  return out;
}

}  // namespace lwo_post
