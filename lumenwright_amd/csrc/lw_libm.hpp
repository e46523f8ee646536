// libm's powf and log2f on the device, equal to the host's for every f32 input of their
// domains here. The reference's CPU code calls them (Rust's f32::powf and f32::log2), and
// glibc's are within one ulp but not always correctly rounded. The device takes the function
// in f64 (error near 2^-52) and rounds it once to f32. Where glibc's result differs from that
// rounding, the exact value lies close to an f32 rounding midpoint, measured exhaustively on
// the CPU: within 2^-33.2 of it relative to the value for powf(x, 1/2.2f) and 2^-32.9 for
// powf(x, 2.2f) over every f32 in [0, 1]; within 2^-29 absolute for log2f over [1, 65536]
// (glibc's log2f errs by up to about 0.75 ulp where its result is small, so its deviations
// keep to an absolute, not a relative, distance). An input whose f64 value lies within the
// function's band of a midpoint -- kLibmBand = 2^-32 relative for powf, kLog2Band = 2^-28
// absolute for log2f -- is "in the band". For those, and only those, the result comes from a
// table the host makes once with libm itself: the sorted pairs (input bits, libm result bits)
// of the band inputs where libm and the device's rounded f64 value differ.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lw {

constexpr double kLibmBand = 0x1p-32;
constexpr double kLog2Band = 0x1p-28;

// Whether the f64 value p, rounded to c, lies in the band: within kLibmBand * |p| of the f32
// rounding midpoint on p's side of c (2^-150, below the least subnormal, when c is 0).
__device__ __forceinline__ bool libm_in_band(double p, float c) {
  const double ap = fabs(p);
  const float ac = fabsf(c);
  if (ac > 0.0f && ac <= 3.4028234e38f) {
    const uint32_t cb = __float_as_uint(ac);
    const float nb = __uint_as_float((double)ac <= ap ? cb + 1u : cb - 1u);
    const double mid = ((double)ac + (double)nb) * 0.5;  // exact: adjacent f32 values
    return fabs(ap - mid) < kLibmBand * ap;
  }
  if (ac == 0.0f && ap > 0.0) return fabs(ap - 0x1p-150) < kLibmBand * ap;
  return false;
}

// The table's value for input bits u, or c when u has no entry.
__device__ __forceinline__ float libm_fixed(uint32_t u, float c, const uint2 *__restrict__ fix, uint32_t n) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (fix[mid].x < u) lo = mid + 1; else hi = mid;
  }
  return (lo < n && fix[lo].x == u) ? __uint_as_float(fix[lo].y) : c;
}

// pow(x, y) in f64 rounded, and whether it is in the band.
__device__ __forceinline__ float pow_band(float x, float y, bool &in_band) {
  const double p = pow((double)x, (double)y);
  const float c = (float)p;
  in_band = libm_in_band(p, c);
  return c;
}

// libm's powf(x, y) for x in [0, 1] (or NaN) with the table for y.
__device__ __forceinline__ float pow_libm(float x, float y, const uint2 *__restrict__ fix, uint32_t nfix) {
  bool in_band;
  const float c = pow_band(x, y, in_band);
  return in_band ? libm_fixed(__float_as_uint(x), c, fix, nfix) : c;
}

// Whether p (> 0), rounded to c (> 0), lies within kLog2Band of the f32 rounding midpoint on
// p's side of c.
__device__ __forceinline__ bool log2_in_band(double p, float c) {
  if (!(c > 0.0f && c <= 3.4028234e38f)) return false;
  const uint32_t cb = __float_as_uint(c);
  const float nb = __uint_as_float((double)c <= p ? cb + 1u : cb - 1u);
  const double mid = ((double)c + (double)nb) * 0.5;
  return fabs(p - mid) < kLog2Band;
}

// log2 in f64 rounded, and whether it is in the band (inputs in (1, 65536] only: the table
// covers those, and log2(1) = 0 is exact).
__device__ __forceinline__ float log2_band(float x, bool &in_band) {
  const double p = log2((double)x);
  const float c = (float)p;
  in_band = x > 1.0f && x <= 65536.0f && log2_in_band(p, c);
  return c;
}

// libm's log2f(x) for x in [1, 65536]; elsewhere the rounded f64 log2.
__device__ __forceinline__ float log2_libm(float x, const uint2 *__restrict__ fix, uint32_t nfix) {
  bool in_band;
  const float c = log2_band(x, in_band);
  return in_band ? libm_fixed(__float_as_uint(x), c, fix, nfix) : c;
}

}  // namespace lw
