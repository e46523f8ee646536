// f32 arithmetic of the HQ path tracer, shared by the gfx950 kernels and the host-side
// frame preparation. Each function follows the reference kernel crate op for op
// (crates/lumenwright-kernel/src/{rng,math}.rs): the same operations in the same order,
// built with -ffp-contract=off, correctly rounded f32 division and square root, and no
// denormal flushing, so that the device rounds every operation as the reference's CPU
// build does. min/max are selects (NaN picks the second operand) and float->int casts
// saturate with NaN -> 0, as Rust's `as` does.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#define LW_HD __host__ __device__ __forceinline__

namespace lw {

// ---- bit casts and Rust `as` casts ----
LW_HD float bits_f(uint32_t u) { return __builtin_bit_cast(float, u); }
LW_HD uint32_t f_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

// Rust's `as` from f32 truncates toward zero, saturates out-of-range values and maps NaN to
// 0. On the device that is exactly what v_cvt_i32_f32 and v_cvt_u32_f32 do (the CDNA ISA:
// "out-of-range floating point values (including infinity) saturate; NaN is converted to 0"),
// one instruction instead of the branches the portable form compiles to.
LW_HD int32_t as_i32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  int32_t r;
  asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
#else
  if (x != x) return 0;
  if (x >= 2147483648.0f) return INT32_MAX;
  if (x < -2147483648.0f) return INT32_MIN;
  return (int32_t)x;
#endif
}
LW_HD uint32_t as_u32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t r;
  asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
#else
  if (x != x) return 0;
  if (x >= 4294967296.0f) return UINT32_MAX;
  if (x <= 0.0f) return 0u;
  return (uint32_t)x;
#endif
}
LW_HD int32_t clamp_i(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---- rng.rs:5-34: PCG-RXS-M-XS hash and the per-(pixel, sample) stream ----
LW_HD uint32_t pcg_hash(uint32_t v) {
  uint32_t state = v * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
  return (word >> 22) ^ word;
}
struct Rng {
  uint32_t state;
  LW_HD static Rng make(uint32_t pixel, uint32_t sample, uint32_t seed) {
    return Rng{pcg_hash(pixel ^ pcg_hash(sample + pcg_hash(seed)))};
  }
  LW_HD float next() {
    state = pcg_hash(state);
    return (float)(state >> 8) * (1.0f / 16777216.0f);
  }
};

// ---- math.rs ----
constexpr float PI_F = 3.14159265358979323846f;
constexpr float FRAC_PI_2_F = 1.57079632679489661923f;
constexpr float SQRT_2_F = 1.41421356237309504880f;
constexpr float LN_2_F = 0.693147180559945309417f;
constexpr float GAMMA = 2.2f;

LW_HD float f_abs(float x) { return x < 0.0f ? -x : x; }
LW_HD float f_min(float a, float b) { return a < b ? a : b; }
LW_HD float f_max(float a, float b) { return a > b ? a : b; }
LW_HD float f_clamp(float x, float lo, float hi) { return f_min(f_max(x, lo), hi); }
LW_HD float f_sqrt(float x) { return sqrtf(x); }
LW_HD float f_floor(float x) { return floorf(x); }

// Rust f32 `%` by 1 and by 2, exactly: a - trunc(a) and a - 2 trunc(a / 2) are exact for
// every finite a (a / 2 is exact, and the difference is a multiple of ulp(a) below the
// divisor); they differ from fmodf only in the sign of a zero result, which no caller
// can observe (each adds 1 to it or takes |. - 1|).
LW_HD float rem1(float a) { return a - truncf(a); }
LW_HD float rem2(float a) { return a - 2.0f * truncf(a * 0.5f); }

// e^x in f32 by the method of glibc's expf (the ARM optimized-routines expf, which Rust
// std's f32::exp reaches through libm on Linux): k = round(x * 32 / ln 2), r the remainder,
// e^x = 2^(k/32) * p(r) with 2^(i/32) from a 32-entry table (each entry less i << 47 so that
// adding k << 47 scales it) and p a cubic, all in f64, rounded once to f32. Checked against
// glibc 2.35's expf over all 2^32 inputs: equal for all but two, whose f64 result lies
// within 1e-15 ulp of an f32 rounding midpoint.
LW_HD float expf_tabulated(float x) {
  const uint32_t ux = f_bits(x);
  const uint32_t abstop = (ux >> 20) & 0x7FFu;
  if (abstop >= 0x42Bu) {  // |x| >= 88
    if (ux == 0xFF800000u) return 0.0f;
    if (abstop >= 0x7F8u) return x + x;
    if (x > 0x1.62e42ep6f) return INFINITY;
    if (x < -0x1.9fe368p6f) return 0.0f;
  }
  constexpr uint64_t T[32] = {
      0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full,
      0x3fef9301d0125b51ull, 0x3fef72b83c7d517bull, 0x3fef54873168b9aaull,
      0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull, 0x3fef06fe0a31b715ull,
      0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
      0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull,
      0x3feea47eb03a5585ull, 0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull,
      0x3feea11473eb0187ull, 0x3feea589994cce13ull, 0x3feeace5422aa0dbull,
      0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
      0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull,
      0x3fef3720dcef9069ull, 0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full,
      0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull};
  constexpr double N = 32.0;
  constexpr double INV_LN2_N = 0x1.71547652b82fep+0 * N;
  constexpr double C0 = 0x1.c6af84b912394p-5 / N / N / N;
  constexpr double C1 = 0x1.ebfce50fac4f3p-3 / N / N;
  constexpr double C2 = 0x1.62e42ff0c52d6p-1 / N;
  constexpr double SHIFT = 0x1.8p+52;
  const double z = INV_LN2_N * (double)x;
  double kd = z + SHIFT;
  const uint64_t ki = __builtin_bit_cast(uint64_t, kd);
  kd -= SHIFT;
  const double r = z - kd;
  const double s = __builtin_bit_cast(double, T[ki % 32] + (ki << 47));
  const double zz = C0 * r + C1;
  const double r2 = r * r;
  double y = C2 * r + 1.0;
  y = zz * r2 + y;
  y = y * s;
  return (float)y;
}

// e^x: libm's expf on the host (what Rust std's f32::exp calls), expf_tabulated on the
// device.
LW_HD float f_exp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return expf_tabulated(x);
#else
  return ::expf(x);
#endif
}

// math.rs:44-64
LW_HD void sin_cos_turns(float u, float &s_out, float &c_out) {
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
  bool neg = (k & 2) != 0;
  s_out = neg ? -s2 : s2;
  c_out = neg ? -c2 : c2;
}

// math.rs:86-103 (Abramowitz & Stegun 4.4.46)
LW_HD float f_acos(float x) {
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
LW_HD float f_atan2(float y, float x) {
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
LW_HD float f_ln(float x) {
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
LW_HD float gamma_to_linear(float c) { return c <= 0.0f ? 0.0f : f_exp(GAMMA * f_ln(c)); }

// math.rs:218-344
struct V3 {
  float x, y, z;
};
LW_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
LW_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
LW_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
LW_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
LW_HD V3 operator/(V3 a, float s) { return a * (1.0f / s); }
LW_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
LW_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LW_HD V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
LW_HD float length(V3 a) { return f_sqrt(dot(a, a)); }
LW_HD V3 normalize(V3 a) {
  float l = length(a);
  return l > 0.0f ? a * (1.0f / l) : a;
}
LW_HD V3 mul(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
LW_HD float max_elem(V3 a) { return f_max(a.x, f_max(a.y, a.z)); }
LW_HD float get(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
LW_HD V3 transmittance(V3 a, float t) {
  return V3{f_exp(-a.x * t), f_exp(-a.y * t), f_exp(-a.z * t)};
}
LW_HD V3 lerp(V3 a, V3 b, float t) { return a + (b - a) * t; }

// math.rs:347-356 (Duff et al. 2017)
LW_HD void onb(V3 n, V3 &t_out, V3 &b_out) {
  float sign = n.z >= 0.0f ? 1.0f : -1.0f;
  float a = -1.0f / (sign + n.z);
  float b = n.x * n.y * a;
  t_out = V3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
  b_out = V3{b, sign + n.y * n.y * a, -n.y};
}

}  // namespace lw
