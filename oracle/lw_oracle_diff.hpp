// ORACLE -- test infrastructure only. The CPU reference of differentiable rendering:
// render_sample (integrator.rs:539-709, as restated in lw_oracle_impl.hpp) with palette,
// intensity and sky parameters, and the gradient of an image loss by path replay. The
// reference has no such feature; this restates the method the device kernels implement
// (radiative backpropagation, Nimier-David et al. 2020; path replay, Vicini et al. 2021)
// on the oracle's own functions, so that tests can check the device's gradients against
// it sample for sample, and check it against central finite differences of its own
// forward, closed forms and the parameters' identity values.
//
// The estimator, per sample, with dl = dL/dI(pixel) / spp:
//   * a scattering vertex with palette group g: dl . (R_final - R_before) / pal_g, where
//     R_before is the radiance before the vertex's emission (the vertex's albedo scales
//     everything the path adds afterwards);
//   * an emission at such a vertex: its second albedo, dl . e / pal_g, and its intensity
//     group, dl . e / inten;
//   * emitter sampling: the colour of the surface it reaches, dl . term / pal, and the
//     sampled face's intensity group, dl . term / inten;
//   * the sky at a miss: dl . beta * dsky_k(d), with dsky_k from the cache's derivative
//     tables through rgb_from_hsl's Jacobian, plus the sky light's derivative;
//   * the sun: the interior term dl . beta A tr cos * dsun_color_k and the rim term of the
//     moving cone, 0.5 sin a / (1 - cos a) * dl . (g(u) - g(-u)) * (dsun_dir_k . u) at a
//     uniform rim direction u drawn from a stream of its own.
// Russian roulette decisions are the forward's; its 1 / q is a constant.
#pragma once

namespace LWO_NS {

static const uint32_t NO_GROUP = 0xFFFFFFFFu;
static const uint32_t MAX_SUFFIX = 16u;

// lw_diff_desc (include/lw/abi.h), byte for byte.
struct DiffDesc {
  uint32_t n_palette, n_intensity;
  const uint32_t *quad_palette;
  const uint32_t *quad_intensity;
  const float *palette;
  const float *intensity;
  uint32_t sky_params, rr_off;
  const float *d_sky_cache;
  float d_sun_dir[3][3];
  float d_sun_color[3][3];
  float d_sky_light[3];
  uint32_t pad;
};
static_assert(sizeof(DiffDesc) == 144, "lw_diff_desc layout");

struct DiffCtx {
  const DiffDesc &dd;
  size_t quads_len;
  size_t cache_len;
  double *grads;  // null in the forward
  uint32_t slots() const { return 3u * dd.n_palette + dd.n_intensity + 3u; }
  uint32_t group(uint32_t q) const {
    if (q >= quads_len) throw OracleError("quad index out of range");
    return dd.quad_palette[q];
  }
  uint32_t igroup(uint32_t q) const {
    if (q >= quads_len) throw OracleError("quad index out of range");
    return dd.quad_intensity[q];
  }
  V3 pal(uint32_t g) const {
    return g == NO_GROUP ? ONE3 : V3{dd.palette[3 * g], dd.palette[3 * g + 1], dd.palette[3 * g + 2]};
  }
  float inten(uint32_t g) const { return g == NO_GROUP ? 1.0f : dd.intensity[g]; }
  void add_pal(uint32_t g, V3 v) const {
    grads[3 * g] += (double)v.x;
    grads[3 * g + 1] += (double)v.y;
    grads[3 * g + 2] += (double)v.z;
  }
  void add_inten(uint32_t g, float v) const { grads[3 * dd.n_palette + g] += (double)v; }
  void add_sky(int k, float v) const { grads[3 * dd.n_palette + dd.n_intensity + k] += (double)v; }
};

static inline V3 vdiv3(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }

// d rgb_from_hsl (integrator.rs:108-127) along the tangent (th, ts, tl), on the branch the
// value takes.
static V3 rgb_from_hsl_tangent(float hue, float sat, float light, float th, float ts, float tl) {
  float a2 = 2.0f * light - 1.0f;
  float one_minus = 1.0f - f_abs(a2);
  float d_one_minus = (a2 < 0.0f ? 2.0f : -2.0f) * tl;
  float cv = one_minus * sat;
  float c = cv, dc = d_one_minus * sat + one_minus * ts;
  if (1.0f < cv) {
    c = 1.0f;
    dc = 0.0f;
  }
  float h = hue * 6.0f, dh = th * 6.0f;
  float r = f_rem(h, 2.0f) - 1.0f;
  float dx = dc * (1.0f - f_abs(r)) - c * (r < 0.0f ? -dh : dh);
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
  float m = light - 0.5f * c;
  float dm = 0.0f > m ? 0.0f : tl - 0.5f * dc;
  return d + v3(dm, dm, dm);
}

// d sky_radiance (integrator.rs:158-188) / d parameter k.
static V3 sky_radiance_tangent(const View &s, const DiffCtx &C, int k, V3 d) {
  const Params &p = *s.params;
  float n = p.misc_f[2];
  if (n < 1.0f) return ZERO3;
  float theta = f_atan2(d.z, d.x) / (2.0f * PI_F);
  theta = f_rem(f_rem(theta, 1.0f) + 1.0f, 1.0f);
  float phi = 1.0f - f_acos(f_clamp(d.y, -1.0f, 1.0f)) / PI_F;
  float x = theta * n, y = phi * n;
  float fx = f_floor(f_clamp(x, 0.0f, n - 1.0f));
  float fy = f_floor(f_clamp(y, 0.0f, n - 1.0f));
  uint32_t m = as_u32(n) + 1u;
  auto idx = [&](float i, float j) -> size_t { return (size_t)((as_u32(i) * m + as_u32(j)) * 3u); };
  size_t a00 = idx(fx, fy), a10 = idx(fx + 1.0f, fy), a01 = idx(fx, fy + 1.0f),
         a11 = idx(fx + 1.0f, fy + 1.0f);
  if (a11 + 2 >= C.cache_len) throw OracleError("sky derivative table too short");
  const float *g = C.dd.d_sky_cache + (size_t)k * C.cache_len;
  float hsl[3], dh[3];
  for (int c = 0; c < 3; c++) {
    float y0 = sky_at(s, a00 + c) + (x - fx) * (sky_at(s, a10 + c) - sky_at(s, a00 + c));
    float y1 = sky_at(s, a01 + c) + (x - fx) * (sky_at(s, a11 + c) - sky_at(s, a01 + c));
    hsl[c] = y0 + (y - fy) * (y1 - y0);
    y0 = g[a00 + c] + (x - fx) * (g[a10 + c] - g[a00 + c]);
    y1 = g[a01 + c] + (x - fx) * (g[a11 + c] - g[a01 + c]);
    dh[c] = y0 + (y - fy) * (y1 - y0);
  }
  V3 drgb = rgb_from_hsl_tangent(hsl[0], hsl[1], hsl[2], dh[0], dh[1], dh[2]);
  return drgb * p.sky_sun[3] + rgb_from_hsl(hsl[0], hsl[1], hsl[2]) * C.dd.d_sky_light[k];
}

// The rim estimate's random number (its own stream, so the path's draws are the forward's).
static inline float rim_draw(uint32_t pixel, uint32_t sample, uint32_t seed, uint32_t bounce) {
  Rng r = Rng::make(pixel, sample, seed ^ (0x5BD1E995u + 0x9E3779B9u * (bounce + 1u)));
  return r.next();
}

// render_sample under the parameters; with C.grads, also the sample's gradient times dl.
static V3 render_sample_diff(const View &s, const DiffCtx &C, uint32_t px, uint32_t py,
                             uint32_t sample, V3 dl) {
  const Params &p = *s.params;
  const bool adj = C.grads != nullptr;
  uint32_t width = p.frame[0];
  uint32_t height = p.frame[1];
  uint32_t max_bounces = p.frame[3];
  float eps = p.misc_f[1];
  Rng rng = Rng::make(py * width + px, sample, p.misc_u[0]);
  uint32_t sfx[MAX_SUFFIX];
  uint32_t nsfx = 0;

  float tan_half = p.cam_pos[3];
  float aspect = p.cam_fwd[3];
  float jx = rng.next();
  float jy = rng.next();
  float sx = (2.0f * ((float)px + jx) / (float)width - 1.0f) * aspect * tan_half;
  float sy = (1.0f - 2.0f * ((float)py + jy) / (float)height) * tan_half;
  V3 d = normalize(from4(p.cam_fwd) + from4(p.cam_right) * sx + from4(p.cam_up) * sy);
  V3 o = from4(p.cam_pos);

  V3 radiance = ZERO3;
  V3 beta = ONE3;
  uint32_t bounces = 0;
  bool specular = true;
  bool in_water = false;
  uint32_t strategy = p.misc_u[3] > 0 ? p.emitter[0] : EMITTER_SAMPLING_NONE;
  float tmin = 0.0f;
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
        if (adj && C.dd.sky_params)
          for (int k = 0; k < 3; k++) C.add_sky(k, dot(dl, mul(beta, sky_radiance_tangent(s, C, k, d))));
        if (specular && bounces == 0) radiance = radiance + mul(beta, sun_texture(s, d));
      }
      break;
    }
    const Hit &h = hit;
    if (in_water) beta = mul(beta, transmittance(sigma, h.t));
    const Quad &q = quad_at(s, h.quad);
    uint32_t flags = q_flags(q);
    V3 pos = o + d * h.t;
    float alpha = (float)(h.texel >> 24) / 255.0f;
    uint32_t g = C.group(h.quad);
    V3 palf = C.pal(g);
    V3 albedo = mul(mul(linear_rgb(h.texel), linear_rgb(q_tint(q))), palf);
    bool front;
    V3 n = shading_normal(q, d, front);

    if ((flags & FLAG_WATER) != 0) {
      water_interface(p, pos, d, n, front, in_water, eps, rng, d, o, in_water);
      specular = true;
      tmin = 0.0f;
      continue;
    }

    if (!forced && (flags & FLAG_ALPHA_MASK) == ALPHA_TRANSLUCENT && rng.next() >= alpha) {
      beta = mul(beta, lerp(ONE3, albedo, alpha));
      o = pos + d * eps;
      tmin = 0.0f;
      continue;
    }

    if (bounces >= max_bounces) break;
    uint32_t ig = C.igroup(h.quad);
    float emittance = q_emittance(q) * p.misc_f[0] * C.inten(ig);
    bool emits = emittance > 0.0f &&
                 (p.emitter[1] == 0 || strategy == EMITTER_SAMPLING_NONE || bounces == 0);
    if (adj && g != NO_GROUP && nsfx < MAX_SUFFIX) {
      sfx[nsfx++] = g;
      C.add_pal(g, -vdiv3(mul(dl, radiance), palf));
    }
    if (emits) {
      V3 e = mul(beta, mul(albedo, albedo)) * emittance;
      radiance = radiance + e;
      if (adj) {
        if (g != NO_GROUP) C.add_pal(g, vdiv3(mul(dl, e), palf));
        if (ig != NO_GROUP) C.add_inten(ig, dot(dl, e) / C.dd.intensity[ig]);
      }
    }
    bounces += 1;

    V3 origin = pos + n * eps;
    V3 sun = from4(p.sun_dir);
    V3 sun_color = from4(p.sun_color);
    float sin_r = f_sqrt(f_max(1.0f - p.sun_dir[3] * p.sun_dir[3], 0.0f));
    bool lit = p.sun_color[0] > 0.0f || p.sun_color[1] > 0.0f || p.sun_color[2] > 0.0f;
    if (lit && dot(n, sun) > -sin_r && p.sun_dir[1] > -sin_r) {
      V3 l = sample_sun(s, rng);
      float cos_l = dot(n, l);
      if (cos_l > 0.0f && l.y >= 0.0f) {
        V3 tr = shadow(s, origin, l, INFINITY);
        radiance = radiance + mul(mul(mul(beta, albedo), sun_color), tr) * cos_l;
        if (adj && C.dd.sky_params) {
          V3 base = mul(mul(beta, albedo), tr) * cos_l;
          for (int k = 0; k < 3; k++) C.add_sky(k, dot(dl, mul(base, v3(C.dd.d_sun_color[k][0], C.dd.d_sun_color[k][1], C.dd.d_sun_color[k][2]))));
        }
      }
      if (adj && C.dd.sky_params) {
        float u = rim_draw(py * width + px, sample, p.misc_u[0], bounces);
        float sin_phi, cos_phi;
        sin_cos_turns(u, sin_phi, cos_phi);
        V3 t, b;
        onb(sun, t, b);
        V3 uv = t * cos_phi + b * sin_phi;
        float cos_a = p.sun_dir[3];
        float wgt = 0.5f * sin_r / (1.0f - cos_a);
        V3 gsum = ZERO3;
        for (int side = 0; side < 2; side++) {
          V3 us = side == 0 ? uv : -uv;
          V3 lr = normalize(sun * cos_a + us * sin_r);
          float cr = dot(n, lr);
          if (cr > 0.0f && lr.y >= 0.0f) {
            V3 trr = shadow(s, origin, lr, INFINITY);
            V3 gr = mul(mul(mul(beta, albedo), sun_color), trr) * cr;
            gsum = side == 0 ? gsum + gr : gsum - gr;
          }
        }
        float dg = dot(dl, gsum) * wgt;
        C.add_sky(1, dg * dot(v3(C.dd.d_sun_dir[1][0], C.dd.d_sun_dir[1][1], C.dd.d_sun_dir[1][2]), uv));
        C.add_sky(2, dg * dot(v3(C.dd.d_sun_dir[2][0], C.dd.d_sun_dir[2][1], C.dd.d_sun_dir[2][2]), uv));
      }
    }
    if (!emits && strategy != EMITTER_SAMPLING_NONE) {
      // sample_emitters for ONE (integrator.rs:441-469 over 418-433 and 380-414).
      uint32_t start, count;
      emitter_cell(s, pos, start, count);
      V3 e = ZERO3;
      uint32_t eg = NO_GROUP, cg = NO_GROUP;
      V3 cpal = ONE3;
      float efac = 1.0f;
      if (count != 0) {
        size_t list = (size_t)EMITTER_GRID_HEADER +
                      (size_t)eg_at(s, 1) * (size_t)eg_at(s, 2) * (size_t)eg_at(s, 3) * 2 + (size_t)start;
        uint32_t ei = eg_at(s, list + pick(count, rng.next()));
        Emitter em = at(s.emitters, s.emitters_len, ei, "emitters");
        uint32_t faces = em.faces[0];
        if (faces != 0) {
          uint32_t f = pick(faces, rng.next());
          uint32_t fq = f_bits(em.pos[3]) + f;
          const Quad &eq = quad_at(s, fq);
          float ss = rng.next();
          float tt = rng.next();
          V3 at_p = from4(em.pos) + from4(eq.p0) + from4(eq.eu) * ss + from4(eq.ev) * tt;
          V3 to = at_p - origin;
          if (!(dot(to, n) <= 0.0f)) {
            float dist = length(to);
            V3 ldir = to / dist;
            Hit hh;
            if (trace(s, origin, ldir, eps, dist + 2.0f * eps, hh) && !(f_abs(hh.t - dist) >= 2.0f * eps)) {
              V3 nq = q_normal(eq);
              float area = length(nq);
              float cl = f_abs(dot(nq, ldir)) / area;
              eg = C.igroup(fq);
              efac = C.inten(eg);
              cg = C.group(hh.quad);
              cpal = C.pal(cg);
              const Quad &hq = quad_at(s, hh.quad);
              V3 col = mul(mul(linear_rgb(hh.texel), linear_rgb(q_tint(hq))), cpal);
              e = col * (cl / f_max(dist * dist, 1.0f) * area * (q_emittance(eq) * efac) * p.misc_f[0] * 1.0f);
            }
          }
        }
        e = e * PI_F;
      }
      V3 term = mul(mul(beta, albedo), e);
      radiance = radiance + term;
      if (adj) {
        if (cg != NO_GROUP) C.add_pal(cg, vdiv3(mul(dl, term), cpal));
        if (eg != NO_GROUP) C.add_inten(eg, dot(dl, term) / efac);
      }
    }

    beta = mul(beta, albedo);
    if (bounces >= 3 && !C.dd.rr_off) {
      float q_rr = f_clamp(max_elem(beta), 0.05f, 0.95f);
      if (rng.next() > q_rr) break;
      beta = beta / q_rr;
    }
    d = cosine_hemisphere(n, rng);
    o = origin;
    tmin = 0.0f;
    specular = false;
  }
  if (adj)
    for (uint32_t k = 0; k < nsfx; k++) C.add_pal(sfx[k], vdiv3(mul(dl, radiance), C.pal(sfx[k])));
  return radiance;
}

static void check_diff(const View &s, const DiffDesc &dd) {
  const Params &p = *s.params;
  uint32_t strategy = p.misc_u[3] > 0 ? p.emitter[0] : EMITTER_SAMPLING_NONE;
  if (strategy != EMITTER_SAMPLING_NONE && strategy != EMITTER_SAMPLING_ONE)
    throw OracleError("emitter sampling NONE or ONE only");
  for (size_t q = 0; q < s.quads_len; q++) {
    if (dd.quad_palette[q] != NO_GROUP && dd.quad_palette[q] >= dd.n_palette) throw OracleError("palette group");
    if (dd.quad_intensity[q] != NO_GROUP && dd.quad_intensity[q] >= dd.n_intensity) throw OracleError("intensity group");
  }
  if (dd.sky_params && (!(p.sky_map[0] < 1.0f) || !(p.misc_f[2] >= 1.0f) || !dd.d_sky_cache))
    throw OracleError("sky parameters need the Preetham cache and its derivative tables");
}

static size_t cache_len_of(const View &s) {
  const Params &p = *s.params;
  if (p.misc_f[2] < 1.0f) return 0;
  size_t n = (size_t)as_u32(p.misc_f[2]) + 1;
  return n * n * 3;
}

// CpuBackend::render_hq (render/cpu/mod.rs:39-87) over render_sample_diff: rows in
// parallel, f64 accumulation of the finite samples, / spp.
// With `dot_weights`, *dot_out gets sum(weights * image) with the image kept in f64 (the
// per-pixel sums / spp before the f32 output), for finite differences.
static void render_diff(const View &view_in, uint32_t spp_in, const DiffDesc &dd, unsigned threads,
                        float *out, const float *dot_weights = nullptr, double *dot_out = nullptr) {
  Params params = *view_in.params;
  params.frame[2] = 0;
  View view = view_in;
  view.params = &params;
  check_diff(view, dd);
  const DiffCtx C{dd, view.quads_len, cache_len_of(view), nullptr};
  const uint32_t w = params.frame[0], h = params.frame[1];
  const uint32_t spp = spp_in > 0 ? spp_in : 1;
  std::atomic<uint32_t> next_row{0};
  std::vector<std::string> errors(threads);
  std::vector<double> row_dot(h, 0.0);
  auto worker = [&](unsigned tid) {
    try {
      for (;;) {
        uint32_t y = next_row.fetch_add(1);
        if (y >= h) break;
        for (uint32_t x = 0; x < w; x++) {
          double acc[3] = {0.0, 0.0, 0.0};
          for (uint32_t sample = 0; sample < spp; sample++) {
            V3 c = render_sample_diff(view, C, x, y, sample, ZERO3);
            if (finite3(c)) {
              acc[0] += (double)c.x;
              acc[1] += (double)c.y;
              acc[2] += (double)c.z;
            }
          }
          const size_t i = ((size_t)y * w + x) * 3;
          for (int c = 0; c < 3; c++) {
            out[i + c] = (float)(acc[c] / (double)spp);
            if (dot_weights) row_dot[y] += (double)dot_weights[i + c] * (acc[c] / (double)spp);
          }
        }
      }
    } catch (const std::exception &e) {
      errors[tid] = e.what();
      next_row.store(h);
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < threads; t++) pool.emplace_back(worker, t);
  worker(0);
  for (auto &t : pool) t.join();
  for (auto &e : errors)
    if (!e.empty()) throw OracleError(e);
  if (dot_out) {
    double s = 0.0;
    for (double r : row_dot) s += r;
    *dot_out = s;
  }
}

// The gradient of sum(dldi * I) for the image render_diff gives: grads gets
// 3 n_palette + n_intensity + 3 doubles.
static void adjoint(const View &view_in, uint32_t spp_in, const DiffDesc &dd, const float *dldi,
                    unsigned threads, double *grads) {
  Params params = *view_in.params;
  params.frame[2] = 0;
  View view = view_in;
  view.params = &params;
  check_diff(view, dd);
  const uint32_t w = params.frame[0], h = params.frame[1];
  const uint32_t spp = spp_in > 0 ? spp_in : 1;
  const float inv_spp = 1.0f / (float)spp;
  const size_t cache_len = cache_len_of(view);
  const uint32_t slots = 3u * dd.n_palette + dd.n_intensity + 3u;
  std::vector<std::vector<double>> part(threads, std::vector<double>(slots, 0.0));
  std::atomic<uint32_t> next_row{0};
  std::vector<std::string> errors(threads);
  auto worker = [&](unsigned tid) {
    try {
      const DiffCtx C{dd, view.quads_len, cache_len, part[tid].data()};
      for (;;) {
        uint32_t y = next_row.fetch_add(1);
        if (y >= h) break;
        for (uint32_t x = 0; x < w; x++) {
          const float *g = dldi + ((size_t)y * w + x) * 3;
          V3 dl = v3(g[0], g[1], g[2]) * inv_spp;
          if (dl.x == 0.0f && dl.y == 0.0f && dl.z == 0.0f) continue;
          for (uint32_t sample = 0; sample < spp; sample++) render_sample_diff(view, C, x, y, sample, dl);
        }
      }
    } catch (const std::exception &e) {
      errors[tid] = e.what();
      next_row.store(h);
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < threads; t++) pool.emplace_back(worker, t);
  worker(0);
  for (auto &t : pool) t.join();
  for (auto &e : errors)
    if (!e.empty()) throw OracleError(e);
  for (uint32_t i = 0; i < slots; i++) {
    double s = 0.0;
    for (unsigned t = 0; t < threads; t++) s += part[t][i];
    grads[i] = s;
  }
}

}  // namespace LWO_NS
