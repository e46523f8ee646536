// Chunky's sky as the kernel crate computes it, shared by the host (lw_host.cpp) and the
// device (the sky-cache and blend kernels), so both evaluate the same operations:
// preetham_radiance (crates/lumenwright-kernel/src/integrator.rs:44-74), rgb_to_hsl
// (integrator.rs:79-104) and one entry of build_sky_cache (integrator.rs:133-148).
// f_exp is libm's expf on the host and its restatement on the device.
#pragma once

#include "lw/abi.h"
#include "lw_math.hpp"

namespace lw {

LW_HD V3 preetham_radiance(const lw_params &p, V3 d) {
  const V3 sun = v3(p.sky_sun[0], p.sky_sun[1], p.sky_sun[2]);
  const float cos_theta = f_max(d.y + p.sky_zenith[3], 0.0f);
  const float cos_gamma = f_clamp(dot(d, sun), -1.0f, 1.0f);
  const float gamma = f_acos(cos_gamma);
  const float cos2_gamma = cos_gamma * cos_gamma;
  float v[3];
  for (int k = 0; k < 3; k++) {
    const float *A = p.perez_a, *B = p.perez_b, *C = p.perez_c, *D = p.perez_d, *E = p.perez_e;
    const float a = cos_theta > 0.0f ? A[k] * f_exp(B[k] / cos_theta) : 0.0f;
    v[k] = (1.0f + a) * (1.0f + C[k] * f_exp(D[k] * gamma) + E[k] * cos2_gamma);
  }
  const float x = p.sky_zenith[0] * v[0];
  const float y = p.sky_zenith[1] * v[1];
  const float lum = p.sky_zenith[2] * v[2];
  if (y <= 5e-8f) return v3(0.0f, 0.0f, 0.0f);
  const float f = lum / y;
  const float cx = x * f, cy = lum, cz = (1.0f - x - y) * f;
  return v3(2.3706743f * cx - 0.9000405f * cy - 0.4706338f * cz,
            -0.513885f * cx + 1.4253036f * cy + 0.0885814f * cz,
            0.0052982f * cx - 0.0146949f * cy + 1.0093968f * cz) *
         0.045f;
}

LW_HD void rgb_to_hsl(V3 c, float out[3]) {
  const float r = f_clamp(c.x, 0.0f, 1.0f), g = f_clamp(c.y, 0.0f, 1.0f), b = f_clamp(c.z, 0.0f, 1.0f);
  const float cmax = f_max(f_max(r, g), b);
  const float cmin = f_min(f_min(r, g), b);
  const float delta = cmax - cmin;
  const float lightness = (cmax + cmin) / 2.0f;
  float hue;
  if (delta == 0.0f) {
    hue = 0.0f;
  } else if (cmax == r) {
    hue = fmodf((g - b) / delta, 6.0f) / 6.0f;  // exact, as Rust's f32 `%`
  } else if (cmax == g) {
    hue = (((b - r) / delta) + 2.0f) / 6.0f;
  } else {
    hue = (((r - g) / delta) + 4.0f) / 6.0f;
  }
  const float saturation = delta < 5e-8f ? 0.0f : delta / (1.0f - f_abs(2.0f * lightness - 1.0f));
  out[0] = hue;
  out[1] = saturation;
  out[2] = lightness;
}

// Entry (i, j) of the sky cache of resolution n: rgb_to_hsl of the Preetham sky towards
// azimuth i / n turns and height j / n * pi - pi / 2.
LW_HD void sky_cache_entry(const lw_params &p, uint32_t n, uint32_t i, uint32_t j, float out[3]) {
  float sin_t, cos_t, sin_p, cos_p;
  sin_cos_turns((float)i / (float)n, sin_t, cos_t);
  // j / n * pi - pi / 2 is j / (2n) - 1/4 turns.
  sin_cos_turns((float)j / (float)(2u * n) - 0.25f, sin_p, cos_p);
  const float r = f_abs(cos_p);
  rgb_to_hsl(preetham_radiance(p, v3(cos_t * r, sin_p, sin_t * r)), out);
}

}  // namespace lw
