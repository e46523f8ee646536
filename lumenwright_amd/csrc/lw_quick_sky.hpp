// QuickSky::background (crates/lumenwright/src/render/sky.rs:403-532) on the device: the
// client's sky along a view direction -- fog colour, sky disc, sunrise glow, sun and moon
// sprites, stars -- with the operations of the reference's f32 code (glam 0.34.1: lerp is
// a * (1 - s) + b * s, clamp is max then min). The sines and cosines it takes of fixed angles
// come from the host (DiscSectors, libm); the two angles it takes of the direction itself
// (the disc's sector, the star bin) are atan2 and asin in f64 rounded once, which only choose
// a sector or a bin.
#pragma once

#include "lw/host.h"
#include "lw_math.hpp"

namespace lw {

// sin, cos of each disc sector's edges th0 = -pi + k * pi/4 and th0 + pi/4 (sky.rs:408-410).
struct DiscSectors {
  float s0[8], c0[8], s1[8], c1[8];
};

constexpr float RIM_SPHERICAL = 512.24994f;  // sky.rs:297

__device__ __forceinline__ float max_r(float a, float b) {  // f32::max
  return isnan(a) ? b : (isnan(b) ? a : (a > b ? a : b));
}
__device__ __forceinline__ float min_r(float a, float b) {  // f32::min
  return isnan(a) ? b : (isnan(b) ? a : (a < b ? a : b));
}
__device__ __forceinline__ V3 min3(V3 a, float b) { return v3(min_r(a.x, b), min_r(a.y, b), min_r(a.z, b)); }
__device__ __forceinline__ V3 lerp3(V3 a, V3 b, float s) { return a * (1.0f - s) + b * s; }
__device__ __forceinline__ V3 xyz(const float *v) { return v3(v[0], v[1], v[2]); }
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
  return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}

// sky.rs:287-295
__device__ __forceinline__ float linear_fog(float v, float start, float end) {
  if (v <= start) return 0.0f;
  if (v >= end) return 1.0f;
  return (v - start) / (end - start);
}

// sky.rs:402-422
__device__ __forceinline__ bool qs_disc(const lw_quick_sky &q, const DiscSectors &sec, V3 d, V3 &out) {
  if (d.y <= 0.0f) return false;
  const V3 p = d * (16.0f / d.y);
  const float step = 0.785398163397448309616f;  // FRAC_PI_4
  const float phi = (float)atan2((double)p.z, (double)p.x);
  const float kf = fminf(fmaxf(floorf((phi + PI_F) / step), 0.0f), 7.0f);
  const int k = (int)kf;
  const float r0x = 512.0f * sec.c0[k], r0z = 512.0f * sec.s0[k];
  const float r1x = 512.0f * sec.c1[k], r1z = 512.0f * sec.s1[k];
  const float det = r0x * r1z - r1x * r0z;
  const float b1 = (p.x * r1z - r1x * p.z) / det;
  const float b2 = (r0x * p.z - p.x * r0z) / det;
  const float s = b1 + b2;
  if (s > 1.0f) return false;
  const float end = q.sky_color[3];
  const float sph = 16.0f * (1.0f - s) + RIM_SPHERICAL * s;
  const float cyl = 16.0f * (1.0f - s) + 512.0f * s;
  const float f = max_r(linear_fog(sph, 0.0f, end), linear_fog(cyl, end, end));
  out = lerp3(xyz(q.sky_color), xyz(q.fog_color), f);
  return true;
}

// sky.rs:425-454: the nearest fan triangle along d.
__device__ __forceinline__ float qs_fan_alpha(const lw_quick_sky &q, V3 d) {
  const V3 c = xyz(q.fan[0]);
  float best_t = INFINITY, alpha = 0.0f;
  for (int k = 0; k < 16; k++) {
    const V3 e1 = xyz(q.fan[k + 1]) - c, e2 = xyz(q.fan[k + 2]) - c;
    const V3 pv = cross3(d, e2);
    const float det = dot(e1, pv);
    if (fabsf(det) < 1e-6f) continue;
    const float inv = 1.0f / det;
    const V3 tv = v3(-c.x, -c.y, -c.z);
    const float u = dot(tv, pv) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) continue;
    const V3 qv = cross3(tv, e1);
    const float v = dot(d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) continue;
    const float t = dot(e2, qv) * inv;
    if (t > 0.0f && t < best_t) {
      best_t = t;
      alpha = (1.0f - u - v) * q.fan[0][3];
    }
  }
  return alpha;
}

// sky.rs:457-483: a sun or moon quad's texel along d (straight alpha). The texel index is
// held inside the atlas (the reference would stop with an index panic past it).
__device__ __forceinline__ float4 qs_body(const float (*b)[4], V3 d, const uint32_t *atlas, uint32_t atlas_w,
                                          uint32_t atlas_len) {
  const float4 none = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (b[0][3] == 0.0f || b[3][2] <= 0.0f) return none;
  const V3 c = xyz(b[0]);
  const float dn = dot(d, c);
  if (dn <= 0.0f) return none;
  const V3 qv = d * (dot(c, c) / dn) - c;
  const float lx = dot(qv, xyz(b[1])), lz = dot(qv, xyz(b[2]));
  if (fabsf(lx) >= 1.0f || fabsf(lz) >= 1.0f) return none;
  const float u = b[3][0] + (lx * 0.5f + 0.5f) * b[3][2];
  const float v = b[3][1] + (lz * 0.5f + 0.5f) * b[3][3];
  const uint64_t at = (uint64_t)as_u32(floorf(v)) * atlas_w + as_u32(floorf(u));
  const uint32_t t = atlas[at < atlas_len ? at : atlas_len - 1];
  return make_float4((float)(t & 0xFFu) / 255.0f, (float)((t >> 8) & 0xFFu) / 255.0f,
                     (float)((t >> 16) & 0xFFu) / 255.0f, (float)(t >> 24) / 255.0f);
}

// sky.rs:486-507 with star_bin (sky.rs:139-150): the stars whose square lies along d.
__device__ __forceinline__ uint32_t qs_star_hits(const lw_quick_sky &q, V3 d, const uint32_t *table) {
  const V3 qd = v3(dot(xyz(q.star_rot[0]), d), dot(xyz(q.star_rot[1]), d), dot(xyz(q.star_rot[2]), d));
  const float az = (float)atan2((double)qd.z, (double)qd.x);
  const float el = (float)asin((double)fminf(fmaxf(qd.y, -1.0f), 1.0f));
  int ia = as_i32(floorf(((az + PI_F) / 6.28318530717958647692f) * (float)LW_STAR_BINS_AZ));
  int ie = as_i32(floorf(((el + 1.57079632679489661923f) / PI_F) * (float)LW_STAR_BINS_EL));
  ia = min(max(ia, 0), (int)LW_STAR_BINS_AZ - 1);
  ie = min(max(ie, 0), (int)LW_STAR_BINS_EL - 1);
  const uint32_t b = (uint32_t)ie * LW_STAR_BINS_AZ + (uint32_t)ia;
  const uint32_t base = LW_STAR_BINS_AZ * LW_STAR_BINS_EL + 1u;
  uint32_t n = 0;
  for (uint32_t e = table[b]; e < table[b + 1]; e++) {
    const float *o = reinterpret_cast<const float *>(table + base + e * LW_STAR_WORDS);
    const V3 c = v3(o[0], o[1], o[2]);
    const float dn = dot(qd, c);
    if (dn <= 0.0f) continue;
    const V3 p = qd * (dot(c, c) / dn) - c;
    if (fabsf(dot(p, v3(o[3], o[4], o[5]))) < 1.0f && fabsf(dot(p, v3(o[6], o[7], o[8]))) < 1.0f) n++;
  }
  return n;
}

// sky.rs:509-532: the sky passes in order, each clamped as the client's 8-bit target.
__device__ __forceinline__ V3 qs_background(const lw_quick_sky &q, const DiscSectors &sec, V3 d,
                                            const uint32_t *atlas, uint32_t atlas_w, uint32_t atlas_len,
                                            const uint32_t *stars) {
  V3 c = xyz(q.fog_color);
  V3 disc;
  if (q.flags[0] > 0.5f && qs_disc(q, sec, d, disc)) c = disc;
  if (q.flags[1] > 0.5f) c = lerp3(c, xyz(q.sunrise), qs_fan_alpha(q, d));
  for (int k = 0; k < 2; k++) {
    const float4 t = qs_body(k == 0 ? q.sun : q.moon, d, atlas, atlas_w, atlas_len);
    c = min3(c + v3(t.x, t.y, t.z) * t.w, 1.0f);
  }
  if (q.flags[2] > 0.5f) {
    const uint32_t n = qs_star_hits(q, d, stars);
    for (uint32_t i = 0; i < n; i++) c = min3(c + v3(q.fog_dist[3], q.fog_dist[3], q.fog_dist[3]), 1.0f);
  }
  return v3(min_r(max_r(c.x, 0.0f), 1.0f), min_r(max_r(c.y, 0.0f), 1.0f), min_r(max_r(c.z, 0.0f), 1.0f));
}

}  // namespace lw
