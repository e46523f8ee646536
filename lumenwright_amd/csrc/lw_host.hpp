// Host-side frame preparation used inside the library (see include/lw/host.h).
#pragma once

#include "lw/host.h"

namespace lw {
// gamma_to_linear(k / 255) for k = 0..255 (math.rs:213-216 with scene.rs:299-306).
void gamma_lut(float out[256]);
// The display encoding's power as the reference's CPU build computes it: libm's powf(x, 1/2.2f)
// (post.rs:141-144, Rust's f32::powf).
float display_pow(float x);
// libm's powf and log2f (Rust's f32::powf and f32::log2).
float libm_powf(float x, float y);
float libm_log2f(float x);
}  // namespace lw
