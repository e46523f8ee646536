// Differentiable rendering: render_sample with scene parameters, and its adjoint by path
// replay (radiative backpropagation: M. Nimier-David, S. Speierer, B. Ruiz, W. Jakob,
// "Radiative Backpropagation", SIGGRAPH 2020; D. Vicini, S. Speierer, W. Jakob, "Path
// Replay Backpropagation", SIGGRAPH 2021).
//
// Parameters (the reference has none; at their identity values the forward below is
// render_sample, integrator.rs:539-709, operation for operation):
//   * palette factors: the albedo of a hit on quad q is the reference's
//     linear_rgb(texel) * linear_rgb(tint) times the RGB factor pal[qgroup[q]], both where
//     a path scatters and where emitter sampling lands (translucent quads, which also
//     filter shadow rays, take no group);
//   * intensity factors: a quad's emittance times misc_f.x is multiplied by inten[qigroup[q]]
//     where a path meets it and where emitter sampling picks it;
//   * the Preetham sky's turbidity and the sun's altitude and azimuth (radians), through
//     tables of the sky cache's derivatives and the derivatives of sun_dir, sun_color and
//     the sky light, all made on the host by central differences of the host sky setup.
//
// The adjoint replays a sample with the same random numbers and operations, so it meets
// the same surfaces and takes the same branches. A factor f that multiplies every
// contribution from some point of the path on (a scattering albedo, through beta) gives
// dL/dI . (R_final - R_before) d(log f); the replay adds the -R_before part there, keeps
// the group, and adds the R_final part when the path ends. Factors of a single contribution (the emission's second
// albedo, the emitter's colour and intensity, the sky's parameters) are settled at once.
// Russian roulette is held fixed (its 1 / q is a constant of the estimator, which stays
// unbiased). Which surface a ray meets does not depend on any parameter.
//
// The sun: sample_sun draws a direction uniformly in the cone around sun_dir (half-angle
// a, cos a = sun_dir.w), so the sun term is C = (1/Omega) * integral over the cone of
// g(l) = V(l) A sun_color max(n.l, 0) [l.y >= 0] dl. Holding the sample fixed in the world
// while the cone moves splits dC/dtheta into an interior term (only sun_color moves) and a
// term on the cone's rim, which moves with normal speed (d sun_dir/dtheta) . u at the rim
// point cos a s + sin a u: dC/dtheta = (1/Omega) * integral over phi of
// g(l(phi)) ((d s/dtheta) . u(phi)) sin a dphi. The replay estimates it with one uniform
// phi and its antithetic phi + pi (two shadow rays on the rim), with weight
// sin a / (1 - cos a). Shadow edges need no separate treatment: a fixed sample's
// visibility does not change with the sun.
#pragma once

#include "lw_device.hpp"

namespace lw {

constexpr uint32_t NO_GROUP = 0xFFFFFFFFu;

struct DiffDev {
  const float *__restrict__ pal;        // 3 * n_pal albedo factors
  const float *__restrict__ inten;      // n_int intensity factors
  const uint32_t *__restrict__ qgroup;  // per quad: palette group or NO_GROUP
  const uint32_t *__restrict__ qigroup; // per quad: intensity group or NO_GROUP
  const float *__restrict__ dsky;       // 3 tables shaped like the sky cache: d/d(turb, alt, az)
  const float *__restrict__ dldi;       // adjoint: dL/dI per pixel (w * h * 3)
  double *__restrict__ grads;           // adjoint: [3 n_pal | n_int | turb, alt, az]
  float dsun[3][3];                     // d sun_dir / d(turb, alt, az)
  float dsun_color[3][3];               // d sun_color / d(turb, alt, az)
  float dsky_light[3];                  // d params.sky_sun[3] / d(turb, alt, az)
  uint32_t n_pal, n_int, rr_off, sky_params;
  uint32_t cache_len;                   // floats per derivative table
  uint32_t pad[3];
};

__device__ __forceinline__ uint32_t diff_slots(const DiffDev &D) { return 3u * D.n_pal + D.n_int + 3u; }

// rgb_from_hsl (integrator.rs:108-127) along a tangent t of (h, s, l): the branch the value
// takes, differentiated. f_min / f_max / f_abs pick their arguments as in the value.
__device__ __forceinline__ V3 rgb_from_hsl_tangent(float hue, float sat, float light, V3 t) {
  const float a2 = 2.0f * light - 1.0f;
  const float one_minus = 1.0f - f_abs(a2);
  const float d_one_minus = (a2 < 0.0f ? 2.0f : -2.0f) * t.z;
  const float cv = one_minus * sat;
  float c = cv, dc = d_one_minus * sat + one_minus * t.y;
  if (1.0f < cv) {
    c = 1.0f;
    dc = 0.0f;
  }
  const float h = hue * 6.0f, dh = t.x * 6.0f;
  const float r = rem2(h) - 1.0f;
  const float dx = dc * (1.0f - f_abs(r)) - c * (r < 0.0f ? -dh : dh);
  V3 d;
  if (h < 1.0f) {
    d = v3(dc, dx, 0.0f);
  } else if (h < 2.0f) {
    d = v3(dx, dc, 0.0f);
  } else if (h < 3.0f) {
    d = v3(0.0f, dc, dx);
  } else if (h < 4.0f) {
    d = v3(0.0f, dx, dc);
  } else if (h < 5.0f) {
    d = v3(dx, 0.0f, dc);
  } else {
    d = v3(dc, 0.0f, dx);
  }
  const float m = light - 0.5f * c;
  const float dm = 0.0f > m ? 0.0f : t.z - 0.5f * dc;
  return d + v3(dm, dm, dm);
}

// d sky_radiance(d) / d(parameter k) (integrator.rs:158-188): the cache's bilinear taps
// over table k, through rgb_from_hsl's Jacobian, plus the sky light's own derivative.
__device__ __forceinline__ V3 sky_radiance_tangent(const DevScene &s, const lw_params &p,
                                                   const DiffDev &D, int k, V3 d) {
  const float n = p.misc_f[2];
  if (n < 1.0f) return V3{0.0f, 0.0f, 0.0f};
  float theta = f_atan2(d.z, d.x) / (2.0f * PI_F);
  theta = rem1(rem1(theta) + 1.0f);
  const float phi = 1.0f - f_acos(f_clamp(d.y, -1.0f, 1.0f)) / PI_F;
  const float x = theta * n, y = phi * n;
  const float fx = f_floor(f_clamp(x, 0.0f, n - 1.0f));
  const float fy = f_floor(f_clamp(y, 0.0f, n - 1.0f));
  const uint32_t m = as_u32(n) + 1u;
  const uint32_t i0 = as_u32(fx), i1 = as_u32(fx + 1.0f), j0 = as_u32(fy), j1 = as_u32(fy + 1.0f);
  const uint32_t a00 = (i0 * m + j0) * 3u, a10 = (i1 * m + j0) * 3u, a01 = (i0 * m + j1) * 3u,
                 a11 = (i1 * m + j1) * 3u;
  const float *g = D.dsky + (size_t)k * D.cache_len;
  float hsl[3], dhsl[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const float *v = s.sky;
    float y0 = v[a00 + c] + (x - fx) * (v[a10 + c] - v[a00 + c]);
    float y1 = v[a01 + c] + (x - fx) * (v[a11 + c] - v[a01 + c]);
    hsl[c] = y0 + (y - fy) * (y1 - y0);
    y0 = g[a00 + c] + (x - fx) * (g[a10 + c] - g[a00 + c]);
    y1 = g[a01 + c] + (x - fx) * (g[a11 + c] - g[a01 + c]);
    dhsl[c] = y0 + (y - fy) * (y1 - y0);
  }
  const V3 drgb = rgb_from_hsl_tangent(hsl[0], hsl[1], hsl[2], v3(dhsl[0], dhsl[1], dhsl[2]));
  return drgb * p.sky_sun[3] + rgb_from_hsl(hsl[0], hsl[1], hsl[2]) * D.dsky_light[k];
}

// Gradient slots in LDS (f64): [3 n_pal palette | n_int intensities | turb, alt, az].
struct GradSink {
  double *lds;
  uint32_t n_pal, n_int;
  __device__ __forceinline__ void pal(uint32_t g, V3 v) const {
    atomicAdd(&lds[3 * g], (double)v.x);
    atomicAdd(&lds[3 * g + 1], (double)v.y);
    atomicAdd(&lds[3 * g + 2], (double)v.z);
  }
  __device__ __forceinline__ void inten(uint32_t g, float v) const {
    atomicAdd(&lds[3 * n_pal + g], (double)v);
  }
  __device__ __forceinline__ void sky(int k, float v) const {
    atomicAdd(&lds[3 * n_pal + n_int + k], (double)v);
  }
};

__device__ __forceinline__ V3 group_factor(const float *pal, uint32_t g) {
  return g == NO_GROUP ? V3{1.0f, 1.0f, 1.0f} : V3{pal[3 * g], pal[3 * g + 1], pal[3 * g + 2]};
}
__device__ __forceinline__ float inten_factor(const float *inten, uint32_t g) {
  return g == NO_GROUP ? 1.0f : inten[g];
}
__device__ __forceinline__ V3 vdiv(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }

// The palette groups of a path's scattering vertices, whose share of the radiance the
// replay settles when the path ends: one per bounce, so the host caps max_bounces at
// kMaxSuffix for the adjoint. Entry k of a lane is at base[k * stride] (LDS on the device).
constexpr uint32_t kMaxSuffix = 16;
struct SuffixList {
  uint32_t *base;
  uint32_t stride;
};

// The random number of the rim estimate: a stream of its own, so that the replay draws
// exactly the forward's numbers from the path's stream.
__device__ __forceinline__ float rim_draw(uint32_t pixel, uint32_t sample, uint32_t seed,
                                          uint32_t bounce) {
  Rng r = Rng::make(pixel, sample, seed ^ (0x5BD1E995u + 0x9E3779B9u * (bounce + 1u)));
  return r.next();
}

// The radiance of (px, py, sample) under the parameters (ADJ false), or its replay, which
// adds dl . dL_sample/dtheta into G (ADJ true; returns the same radiance). Emitter sampling
// strategies NONE and ONE (the host checks).
// s_in and p_in are the kernel's own first two arguments; after each event's trace the event
// reads them again from the kernel argument segment (kernel_args, lw_device.hpp).
template <bool ADJ, class Cam>
__device__ __forceinline__ V3 render_sample_diff(const DevScene &s_in, const lw_params &p_in,
                                                 const Cam &c, const float *lut,
                                                 const DiffDev &D, uint32_t px, uint32_t py,
                                                 uint32_t sample, const GradSink &G,
                                                 const SuffixList &sfx, V3 dl) {
  const uint32_t width = cam_width(c);
  const uint32_t height = cam_height(c);
  Rng rng = Rng::make(py * width + px, sample, cam_seed(c));
  uint32_t nsfx = 0;

  const float tan_half = c.cam_pos[3];
  const float aspect = c.cam_fwd[3];
  const float jx = rng.next();
  const float jy = rng.next();
  const float sx = (2.0f * ((float)px + jx) / (float)width - 1.0f) * aspect * tan_half;
  const float sy = (1.0f - 2.0f * ((float)py + jy) / (float)height) * tan_half;
  V3 d = normalize(V3{c.cam_fwd[0], c.cam_fwd[1], c.cam_fwd[2]} +
                   V3{c.cam_right[0], c.cam_right[1], c.cam_right[2]} * sx +
                   V3{c.cam_up[0], c.cam_up[1], c.cam_up[2]} * sy);
  V3 o = V3{c.cam_pos[0], c.cam_pos[1], c.cam_pos[2]};

  V3 radiance{0.0f, 0.0f, 0.0f};
  V3 beta{1.0f, 1.0f, 1.0f};
  uint32_t bounces = 0;
  bool specular = true;
  bool in_water = false;
  float tmin = 0.0f;

  for (uint32_t ev = 0; ev < MAX_EVENTS; ev++) {
    Hit hit;
    bool has_hit = trace(s_in, p_in, o, d, tmin, INFINITY, hit, nullptr);
    const KernelArgs &ka = kernel_args();
    const DevScene &s = ka.s;
    const lw_params &p = ka.p;
    const uint32_t max_bounces = p.frame[3];
    const float eps = p.misc_f[1];
    const uint32_t strategy = p.misc_u[3] > 0 ? p.emitter[0] : LW_EMITTER_SAMPLING_NONE;
    const V3 sigma = V3{p.water_color[0], p.water_color[1], p.water_color[2]};
    const V3 sun = V3{p.sun_dir[0], p.sun_dir[1], p.sun_dir[2]};
    const V3 sun_color = V3{p.sun_color[0], p.sun_color[1], p.sun_color[2]};
    const float sin_r = f_sqrt(f_max(1.0f - p.sun_dir[3] * p.sun_dir[3], 0.0f));
    const bool lit = p.sun_color[0] > 0.0f || p.sun_color[1] > 0.0f || p.sun_color[2] > 0.0f;
    bool forced = false;
    if (bounces == 0 && p.see_through[1] > 0 && has_hit && !is_see_through(p, hit.quad)) {
      Hit st;
      if (see_through_hit(s, p, o, d, hit.t, st)) {
        const uint32_t sflags = f_bits(s.quads[(size_t)st.quad * 5].w);
        if ((sflags & LW_FLAG_ALPHA_MASK) == LW_ALPHA_TRANSLUCENT) {
          if (rng.next() < (float)(st.texel >> 24) / 255.0f) {
            hit = st;
            forced = true;
          }
        } else {
          hit = st;
        }
      }
    }
    float tp;
    if (water_plane(p, o, d, tp) && (!has_hit || tp < hit.t)) {
      if (in_water) beta = mul(beta, transmittance(sigma, tp));
      const bool front = d.y < 0.0f;
      const V3 n = v3(0.0f, front ? 1.0f : -1.0f, 0.0f);
      water_interface(p, o + d * tp, d, n, front, in_water, eps, rng, d, o, in_water);
      specular = true;
      tmin = 0.0f;
      continue;
    }
    if (!has_hit) {
      if (!in_water && !under_water_plane(p, o, d)) {
        if (bounces == 0) {
          radiance = radiance + mul(beta, sky_lens(s, p, d));
        } else {
          radiance = radiance + mul(beta, sky_light(s, p, d));
        }
        if (ADJ && D.sky_params) {
#pragma unroll
          for (int k = 0; k < 3; k++) G.sky(k, dot(dl, mul(beta, sky_radiance_tangent(s, p, D, k, d))));
        }
        if (specular && bounces == 0) radiance = radiance + mul(beta, sun_texture(s, p, lut, d));
      }
      break;
    }
    if (in_water) beta = mul(beta, transmittance(sigma, hit.t));
    const float4 *q = s.quads + (size_t)hit.quad * 5;
    const float4 q0 = q[0], q1 = q[1], q2 = q[2];
    const uint32_t flags = f_bits(q0.w);
    const V3 pos = o + d * hit.t;
    const float alpha = (float)(hit.texel >> 24) / 255.0f;
    const uint32_t g = D.qgroup[hit.quad];
    const V3 palf = group_factor(D.pal, g);
    const V3 albedo = mul(mul(linear_rgb(lut, hit.texel), linear_rgb(lut, f_bits(q1.w))), palf);
    // integrator.rs:307-315 (shading_normal)
    V3 n = normalize(cross(xyz(q2), xyz(q1)));
    const bool front = dot(n, d) < 0.0f;
    if (!front) n = -n;

    if ((flags & LW_FLAG_WATER) != 0) {
      water_interface(p, pos, d, n, front, in_water, eps, rng, d, o, in_water);
      specular = true;
      tmin = 0.0f;
      continue;
    }

    if (!forced && (flags & LW_FLAG_ALPHA_MASK) == LW_ALPHA_TRANSLUCENT && rng.next() >= alpha) {
      beta = mul(beta, lerp(V3{1.0f, 1.0f, 1.0f}, albedo, alpha));
      o = pos + d * eps;
      tmin = 0.0f;
      continue;
    }

    if (bounces >= max_bounces) break;
    const uint32_t ig = D.qigroup[hit.quad];
    const float emittance = q2.w * p.misc_f[0] * inten_factor(D.inten, ig);
    const bool emits = emittance > 0.0f &&
                       (p.emitter[1] == 0 || strategy == LW_EMITTER_SAMPLING_NONE || bounces == 0);
    // This vertex's albedo scales everything the path adds from here on (through beta),
    // its emission included: its group gets dl . (R_final - R_now) / pal, the second term
    // now and the first when the path ends. The emission's second albedo is settled below.
    if (ADJ && g != NO_GROUP && nsfx < kMaxSuffix) {
      sfx.base[nsfx * sfx.stride] = g;
      nsfx++;
      G.pal(g, -vdiv(mul(dl, radiance), palf));
    }
    if (emits) {
      const V3 e = mul(beta, mul(albedo, albedo)) * emittance;
      radiance = radiance + e;
      if (ADJ) {
        if (g != NO_GROUP) G.pal(g, vdiv(mul(dl, e), palf));
        if (ig != NO_GROUP) G.inten(ig, dot(dl, e) / D.inten[ig]);
      }
    }
    bounces += 1;

    const V3 origin = pos + n * eps;
    if (lit && dot(n, sun) > -sin_r && p.sun_dir[1] > -sin_r) {
      const V3 l = sample_sun(p, rng);
      const float cos_l = dot(n, l);
      if (cos_l > 0.0f && l.y >= 0.0f) {
        const V3 tr = shadow(s, p, lut, origin, l, INFINITY);
        radiance = radiance + mul(mul(mul(beta, albedo), sun_color), tr) * cos_l;
        if (ADJ && D.sky_params) {
          const V3 base = mul(mul(beta, albedo), tr) * cos_l;
#pragma unroll
          for (int k = 0; k < 3; k++)
            G.sky(k, dot(dl, mul(base, V3{D.dsun_color[k][0], D.dsun_color[k][1], D.dsun_color[k][2]})));
        }
      }
      if (ADJ && D.sky_params) {
        // The rim of the sun's cone, at phi and phi + pi.
        const float u = rim_draw(py * width + px, sample, cam_seed(c), bounces);
        float sin_phi, cos_phi;
        sin_cos_turns(u, sin_phi, cos_phi);
        V3 t, b;
        onb(sun, t, b);
        const V3 uv = t * cos_phi + b * sin_phi;
        const float cos_a = p.sun_dir[3];
        const float wgt = 0.5f * sin_r / (1.0f - cos_a);
        V3 gsum{0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int side = 0; side < 2; side++) {
          const V3 us = side == 0 ? uv : -uv;
          const V3 lr = normalize(sun * cos_a + us * sin_r);
          const float cr = dot(n, lr);
          if (cr > 0.0f && lr.y >= 0.0f) {
            const V3 trr = shadow(s, p, lut, origin, lr, INFINITY);
            const V3 gr = mul(mul(mul(beta, albedo), sun_color), trr) * cr;
            gsum = side == 0 ? gsum + gr : gsum - gr;
          }
        }
        const float dg = dot(dl, gsum) * wgt;
        G.sky(1, dg * dot(V3{D.dsun[1][0], D.dsun[1][1], D.dsun[1][2]}, uv));
        G.sky(2, dg * dot(V3{D.dsun[2][0], D.dsun[2][1], D.dsun[2][2]}, uv));
      }
    }
    if (!emits && strategy != LW_EMITTER_SAMPLING_NONE) {
      // sample_emitters for ONE (integrator.rs:441-469 over 418-433 and 380-414).
      uint32_t start, count;
      emitter_cell(s, pos, start, count);
      V3 e{0.0f, 0.0f, 0.0f};
      uint32_t eg = NO_GROUP, cg = NO_GROUP;
      V3 cpal{1.0f, 1.0f, 1.0f};
      float efac = 1.0f;
      if (count != 0) {
        const uint32_t *gr = s.egrid;
        const size_t list = (size_t)LW_EMITTER_GRID_HEADER + (size_t)gr[1] * gr[2] * gr[3] * 2 + start;
        const lw_emitter em = s.emitters[gr[list + pick(count, rng.next())]];
        const uint32_t faces = em.faces[0];
        if (faces != 0) {
          const uint32_t f = pick(faces, rng.next());
          const uint32_t fq = f_bits(em.pos[3]) + f;
          const float4 *fqp = s.quads + (size_t)fq * 5;
          const float ss = rng.next();
          const float tt = rng.next();
          const float4 e0 = fqp[0], e1 = fqp[1], e2 = fqp[2];
          const V3 at = V3{em.pos[0], em.pos[1], em.pos[2]} + xyz(e0) + xyz(e1) * ss + xyz(e2) * tt;
          const V3 to = at - origin;
          if (!(dot(to, n) <= 0.0f)) {
            const float dist = length(to);
            const V3 ldir = to / dist;
            Hit h;
            if (trace(s, p, origin, ldir, eps, dist + 2.0f * eps, h, nullptr) &&
                !(f_abs(h.t - dist) >= 2.0f * eps)) {
              const V3 nq = cross(xyz(e2), xyz(e1));
              const float area = length(nq);
              const float cl = f_abs(dot(nq, ldir)) / area;
              eg = D.qigroup[fq];
              efac = inten_factor(D.inten, eg);
              cg = D.qgroup[h.quad];
              cpal = group_factor(D.pal, cg);
              const float4 *hq = s.quads + (size_t)h.quad * 5;
              const V3 col = mul(mul(linear_rgb(lut, h.texel), linear_rgb(lut, f_bits(hq[1].w))), cpal);
              e = col * (cl / f_max(dist * dist, 1.0f) * area * (e2.w * efac) * p.misc_f[0] * 1.0f);
            }
          }
        }
        e = e * PI_F;
      }
      const V3 term = mul(mul(beta, albedo), e);
      radiance = radiance + term;
      if (ADJ) {
        if (cg != NO_GROUP) G.pal(cg, vdiv(mul(dl, term), cpal));
        if (eg != NO_GROUP) G.inten(eg, dot(dl, term) / efac);
      }
    }

    beta = mul(beta, albedo);
    if (bounces >= 3 && !D.rr_off) {
      const float q_rr = f_clamp(max_elem(beta), 0.05f, 0.95f);
      if (rng.next() > q_rr) break;
      beta = beta / q_rr;
    }
    d = cosine_hemisphere(n, rng);
    o = origin;
    tmin = 0.0f;
    specular = false;
  }
  if (ADJ) {
    for (uint32_t k = 0; k < nsfx; k++) {
      const uint32_t gk = sfx.base[k * sfx.stride];
      G.pal(gk, vdiv(mul(dl, radiance), group_factor(D.pal, gk)));
    }
  }
  return radiance;
}

}  // namespace lw
