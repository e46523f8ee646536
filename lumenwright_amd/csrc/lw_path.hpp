// render_sample (crates/lumenwright-kernel/src/integrator.rs:539-709) as a per-lane state
// machine with a single trace site.
//
// A sample issues three kinds of ray: the path's own extension rays, the sun's shadow
// segments (integrator.rs:318-346) and emitter face rays (integrator.rs:380-414). The
// reference traces them from three nested call sites, so the lanes of a wave that need a
// shadow ray run it while the others wait. Here every lane holds its path as state and
// names the next ray it needs; the wave then runs one trace for all lanes, whatever kind
// of ray each lane traces, and each lane consumes its result. No trace consumes random
// numbers, so the RNG draws happen in the reference's order (sun cone, emitter pick and
// face point, Russian roulette, bounce direction), and each contribution is added to the
// radiance in the reference's order with the reference's operands.
#pragma once

#include "lw_device.hpp"

namespace lw {

enum : uint32_t { ST_IDLE = 0, ST_EXTEND = 1, ST_SHADOW = 2, ST_EMITTER = 3, ST_DONE = 4 };

struct PathState {
  uint32_t stage;
  uint32_t item;  // the work item the path renders
  Rng rng;
  V3 ro, rd;      // the pending ray (for ST_EXTEND, the path's own o and d)
  float rtmin, rtmax;
  V3 beta, radiance;
  uint32_t bounces, events;
  bool specular, in_water;
  // The bounce in progress while its sun and emitter rays are pending.
  V3 ba;      // beta * albedo
  V3 n;       // shading normal
  V3 origin;  // pos + n * eps, where the bounce's rays leave from
  bool do_emitters;
  V3 tr;      // sun shadow transmittance
  float cos_l;
  uint32_t seg;
  // Emitter sampling in progress.
  uint32_t em_list, em_count, em_k, em_f, em_faces, em_index;
  V3 em_inner, em_outer;
  float em_scale, em_dist;
};

// Built from the kernel's own (p, s); its methods read them again from the kernel argument
// segment where they use them (kernel_args, lw_device.hpp), which kept the persistent form from
// holding every field in spilled scalar registers: 793 -> 954 Msamples/s at the 720p frame.
struct Tracer {
  const DevScene &s;
  const lw_params &p;
  const float *lut;

  __device__ __forceinline__ V3 sun_color() const {
    const KernelArgs &ka = kernel_args();
    const lw_params &p = ka.p;
    return V3{p.sun_color[0], p.sun_color[1], p.sun_color[2]};
  }

  // The camera ray of (px, py, sample) (integrator.rs:540-569).
  template <class Cam>
  __device__ __forceinline__ void start(PathState &P, const Cam &c, uint32_t px, uint32_t py,
                                        uint32_t sample) const {
    const uint32_t width = cam_width(c), height = cam_height(c);
    P.rng = Rng::make(py * width + px, sample, cam_seed(c));
    const float tan_half = c.cam_pos[3];
    const float aspect = c.cam_fwd[3];
    const float jx = P.rng.next();
    const float jy = P.rng.next();
    const float sx = (2.0f * ((float)px + jx) / (float)width - 1.0f) * aspect * tan_half;
    const float sy = (1.0f - 2.0f * ((float)py + jy) / (float)height) * tan_half;
    P.rd = normalize(V3{c.cam_fwd[0], c.cam_fwd[1], c.cam_fwd[2]} +
                     V3{c.cam_right[0], c.cam_right[1], c.cam_right[2]} * sx +
                     V3{c.cam_up[0], c.cam_up[1], c.cam_up[2]} * sy);
    P.ro = V3{c.cam_pos[0], c.cam_pos[1], c.cam_pos[2]};
    P.radiance = V3{0.0f, 0.0f, 0.0f};
    P.beta = V3{1.0f, 1.0f, 1.0f};
    P.bounces = 0;
    P.events = 0;
    P.specular = true;
    P.in_water = false;
    P.rtmin = 0.0f;
    extend(P);
  }

  // The next iteration of the event loop (integrator.rs:571): its trace, or the end of
  // the path after MAX_EVENTS iterations.
  __device__ __forceinline__ void extend(PathState &P) const {
    if (P.events >= MAX_EVENTS) {
      P.stage = ST_IDLE;
      return;
    }
    P.events++;
    P.rtmax = INFINITY;
    P.stage = ST_EXTEND;
  }

  __device__ __forceinline__ uint32_t strategy() const {
    const KernelArgs &ka = kernel_args();
    const lw_params &p = ka.p;
    return p.misc_u[3] > 0 ? p.emitter[0] : LW_EMITTER_SAMPLING_NONE;
  }

  // Consumes the result of the pending ray.
  __device__ __forceinline__ void consume(PathState &P, bool has_hit, const Hit &h) const {
    if (P.stage == ST_EXTEND)
      on_extend(P, has_hit, h);
    else if (P.stage == ST_SHADOW)
      on_shadow(P, has_hit, h);
    else
      on_emitter(P, has_hit, h);
  }

  // integrator.rs:572-676: one event of the path, up to the sun's cone sample.
  __device__ __forceinline__ void on_extend(PathState &P, bool has_hit, Hit hit) const {
    const KernelArgs &ka = kernel_args();
    const DevScene &s = ka.s;
    const lw_params &p = ka.p;
    const float eps = p.misc_f[1];
    V3 o = P.ro, d = P.rd;
    bool forced = false;
    if (P.bounces == 0 && p.see_through[1] > 0 && has_hit && !is_see_through(p, hit.quad)) {
      Hit st;
      if (see_through_hit(s, p, o, d, hit.t, st)) {
        const uint32_t sflags = f_bits(s.quads[(size_t)st.quad * 5].w);
        if ((sflags & LW_FLAG_ALPHA_MASK) == LW_ALPHA_TRANSLUCENT) {
          if (P.rng.next() < (float)(st.texel >> 24) / 255.0f) {
            hit = st;
            forced = true;
          }
        } else {
          hit = st;
        }
      }
    }
    const V3 sigma = V3{p.water_color[0], p.water_color[1], p.water_color[2]};
    float tp;
    if (water_plane(p, o, d, tp) && (!has_hit || tp < hit.t)) {
      if (P.in_water) P.beta = mul(P.beta, transmittance(sigma, tp));
      const bool front = d.y < 0.0f;
      const V3 n = v3(0.0f, front ? 1.0f : -1.0f, 0.0f);
      water_interface(p, o + d * tp, d, n, front, P.in_water, eps, P.rng, P.rd, P.ro, P.in_water);
      P.specular = true;
      P.rtmin = 0.0f;
      extend(P);
      return;
    }
    if (!has_hit) {
      if (!P.in_water && !under_water_plane(p, o, d)) {
        if (P.bounces == 0) {
          P.radiance = P.radiance + mul(P.beta, sky_lens(s, p, d));
        } else {
          P.radiance = P.radiance + mul(P.beta, sky_light(s, p, d));
        }
        if (P.specular && P.bounces == 0)
          P.radiance = P.radiance + mul(P.beta, sun_texture(s, p, lut, d));
      }
      P.stage = ST_IDLE;
      return;
    }
    if (P.in_water) P.beta = mul(P.beta, transmittance(sigma, hit.t));
    const float4 *q = s.quads + (size_t)hit.quad * 5;
    const float4 q0 = q[0], q1 = q[1], q2 = q[2];
    const uint32_t flags = f_bits(q0.w);
    const V3 pos = o + d * hit.t;
    const float alpha = (float)(hit.texel >> 24) / 255.0f;
    const V3 albedo = mul(linear_rgb(lut, hit.texel), linear_rgb(lut, f_bits(q1.w)));
    // integrator.rs:307-315 (shading_normal)
    V3 n = normalize(cross(xyz(q2), xyz(q1)));
    const bool front = dot(n, d) < 0.0f;
    if (!front) n = -n;

    if ((flags & LW_FLAG_WATER) != 0) {
      water_interface(p, pos, d, n, front, P.in_water, eps, P.rng, P.rd, P.ro, P.in_water);
      P.specular = true;
      P.rtmin = 0.0f;
      extend(P);
      return;
    }
    if (!forced && (flags & LW_FLAG_ALPHA_MASK) == LW_ALPHA_TRANSLUCENT && P.rng.next() >= alpha) {
      P.beta = mul(P.beta, lerp(V3{1.0f, 1.0f, 1.0f}, albedo, alpha));
      P.ro = pos + d * eps;
      P.rtmin = 0.0f;
      extend(P);
      return;
    }
    if (P.bounces >= p.frame[3]) {
      P.stage = ST_IDLE;
      return;
    }
    const uint32_t strat = strategy();
    const float emittance = q2.w * p.misc_f[0];
    const bool emits = emittance > 0.0f &&
                       (p.emitter[1] == 0 || strat == LW_EMITTER_SAMPLING_NONE || P.bounces == 0);
    if (emits) P.radiance = P.radiance + mul(P.beta, mul(albedo, albedo)) * emittance;
    P.bounces += 1;

    P.origin = pos + n * eps;
    P.n = n;
    P.ba = mul(P.beta, albedo);
    P.do_emitters = !emits && strat != LW_EMITTER_SAMPLING_NONE;
    if (P.do_emitters) emitter_cell(s, pos, P.em_list, P.em_count);

    const V3 sun = V3{p.sun_dir[0], p.sun_dir[1], p.sun_dir[2]};
    const float sin_r = f_sqrt(f_max(1.0f - p.sun_dir[3] * p.sun_dir[3], 0.0f));
    const bool lit = p.sun_color[0] > 0.0f || p.sun_color[1] > 0.0f || p.sun_color[2] > 0.0f;
    if (lit && dot(n, sun) > -sin_r && p.sun_dir[1] > -sin_r) {
      const V3 l = sample_sun(p, P.rng);
      const float cos_l = dot(n, l);
      if (cos_l > 0.0f && l.y >= 0.0f) {
        P.tr = V3{1.0f, 1.0f, 1.0f};
        P.cos_l = cos_l;
        P.seg = 0;
        P.ro = P.origin;
        P.rd = l;
        P.rtmin = eps;
        P.rtmax = INFINITY;
        P.stage = ST_SHADOW;
        return;
      }
    }
    after_sun(P);
  }

  // integrator.rs:322-345: one segment of the sun's shadow ray.
  __device__ __forceinline__ void on_shadow(PathState &P, bool has_hit, const Hit &h) const {
    const KernelArgs &ka = kernel_args();
    const DevScene &s = ka.s;
    const lw_params &p = ka.p;
    const float eps = p.misc_f[1];
    if (has_hit) {
      const float4 *q = s.quads + (size_t)h.quad * 5;
      const float4 q0 = q[0], q1 = q[1], q2 = q[2];
      const uint32_t flags = f_bits(q0.w);
      bool blocked = false;
      if ((flags & LW_FLAG_WATER) != 0) {
        if (dot(cross(xyz(q2), xyz(q1)), P.rd) > 0.0f)
          P.tr = mul(P.tr, transmittance(V3{p.water_color[0], p.water_color[1], p.water_color[2]}, h.t));
      } else if ((flags & LW_FLAG_ALPHA_MASK) == LW_ALPHA_TRANSLUCENT) {
        const float a = (float)(h.texel >> 24) / 255.0f;
        const V3 col = mul(linear_rgb(lut, h.texel), linear_rgb(lut, f_bits(q1.w)));
        P.tr = mul(P.tr, lerp(V3{1.0f, 1.0f, 1.0f}, col, a));
      } else {
        P.tr = V3{0.0f, 0.0f, 0.0f};
        blocked = true;
      }
      if (!blocked) {
        P.ro = P.ro + P.rd * (h.t + eps);
        P.rtmax -= h.t + eps;
        P.seg += 1;
        if (P.seg < 8) return;  // the next segment is pending
      }
    }
    // integrator.rs:685
    P.radiance = P.radiance + mul(mul(P.ba, sun_color()), P.tr) * P.cos_l;
    after_sun(P);
  }

  // integrator.rs:688-706: emitter sampling, then the bounce.
  __device__ __forceinline__ void after_sun(PathState &P) const {
    if (P.do_emitters) {
      begin_emitters(P);
      return;
    }
    finish_bounce(P);
  }

  // integrator.rs:695-706
  __device__ __forceinline__ void finish_bounce(PathState &P) const {
    P.beta = P.ba;
    if (P.bounces >= 3) {
      const float q_rr = f_clamp(max_elem(P.beta), 0.05f, 0.95f);
      if (P.rng.next() > q_rr) {
        P.stage = ST_IDLE;
        return;
      }
      P.beta = P.beta / q_rr;
    }
    P.rd = cosine_hemisphere(P.n, P.rng);
    P.ro = P.origin;
    P.rtmin = 0.0f;
    P.specular = false;
    extend(P);
  }

  // integrator.rs:441-469 and 418-433 up to the first face ray, or to the end.
  __device__ __forceinline__ void begin_emitters(PathState &P) const {
    const KernelArgs &ka = kernel_args();
    const DevScene &s = ka.s;
    const lw_params &p = ka.p;
    if (P.em_count == 0) {
      add_emitters(P, V3{0.0f, 0.0f, 0.0f});
      return;
    }
    const uint32_t *g = s.egrid;
    P.em_list = LW_EMITTER_GRID_HEADER + g[1] * g[2] * g[3] * 2u + P.em_list;
    const uint32_t strat = p.emitter[0];
    if (strat == LW_EMITTER_SAMPLING_ONE || strat == LW_EMITTER_SAMPLING_ONE_BLOCK) {
      P.em_index = g[P.em_list + pick(P.em_count, P.rng.next())];
      P.em_faces = s.emitters[P.em_index].faces[0];
      if (P.em_faces == 0) {
        add_emitters(P, V3{0.0f, 0.0f, 0.0f} * PI_F);
        return;
      }
      if (strat == LW_EMITTER_SAMPLING_ONE) {
        P.em_f = pick(P.em_faces, P.rng.next());
      } else {
        P.em_f = 0;
        P.em_inner = V3{0.0f, 0.0f, 0.0f};
      }
      next_face(P);
      return;
    }
    if (strat == LW_EMITTER_SAMPLING_ALL) {
      P.em_outer = V3{0.0f, 0.0f, 0.0f};
      P.em_k = 0;
      next_emitter(P);
      return;
    }
    add_emitters(P, V3{0.0f, 0.0f, 0.0f});
  }

  // ALL: the next emitter of the cell with faces, or the end of the sum.
  __device__ __forceinline__ void next_emitter(PathState &P) const {
    const KernelArgs &ka = kernel_args();
    const DevScene &s = ka.s;
    while (P.em_k < P.em_count) {
      P.em_index = s.egrid[P.em_list + P.em_k];
      P.em_faces = s.emitters[P.em_index].faces[0];
      if (P.em_faces != 0) {
        P.em_f = 0;
        P.em_inner = V3{0.0f, 0.0f, 0.0f};
        next_face(P);
        return;
      }
      P.em_outer = P.em_outer + V3{0.0f, 0.0f, 0.0f};  // sample_emitter of no faces
      P.em_k += 1;
    }
    add_emitters(P, P.em_outer * (PI_F / (float)P.em_count));
  }

  // integrator.rs:380-401: the face's point and ray; faces behind the surface give zero.
  __device__ __forceinline__ void next_face(PathState &P) const {
    const KernelArgs &ka = kernel_args();
    const DevScene &s = ka.s;
    const lw_params &p = ka.p;
    const uint32_t strat = p.emitter[0];
    for (;;) {
      const lw_emitter e = s.emitters[P.em_index];
      const float scaler = strat == LW_EMITTER_SAMPLING_ONE ? 1.0f : 1.0f / (float)P.em_faces;
      const float4 *q = s.quads + (size_t)(f_bits(e.pos[3]) + P.em_f) * 5;
      const float ss = P.rng.next();
      const float t = P.rng.next();
      const float4 q0 = q[0], q1 = q[1], q2 = q[2];
      const V3 at = V3{e.pos[0], e.pos[1], e.pos[2]} + xyz(q0) + xyz(q1) * ss + xyz(q2) * t;
      const V3 to = at - P.origin;
      if (!(dot(to, P.n) <= 0.0f)) {
        const float dist = length(to);
        const V3 l = to / dist;
        const float eps = p.misc_f[1];
        const V3 nq = cross(xyz(q2), xyz(q1));
        const float area = length(nq);
        const float cos_l = f_abs(dot(nq, l)) / area;
        P.em_scale = cos_l / f_max(dist * dist, 1.0f) * area * q2.w * p.misc_f[0] * scaler;
        P.em_dist = dist;
        P.ro = P.origin;
        P.rd = l;
        P.rtmin = eps;
        P.rtmax = dist + 2.0f * eps;
        P.stage = ST_EMITTER;
        return;
      }
      if (!face_done(P, V3{0.0f, 0.0f, 0.0f})) return;
    }
  }

  // A face's value is known: fold it into the sum. Returns true when another face of the
  // same emitter follows (the caller sets it up), false when the sampling moved on.
  __device__ __forceinline__ bool face_done(PathState &P, V3 value) const {
    const KernelArgs &ka = kernel_args();
    const lw_params &p = ka.p;
    const uint32_t strat = p.emitter[0];
    if (strat == LW_EMITTER_SAMPLING_ONE) {
      add_emitters(P, value * PI_F);
      return false;
    }
    P.em_inner = P.em_inner + value;
    P.em_f += 1;
    if (P.em_f < P.em_faces) return true;
    if (strat == LW_EMITTER_SAMPLING_ONE_BLOCK) {
      add_emitters(P, P.em_inner * PI_F);
      return false;
    }
    P.em_outer = P.em_outer + P.em_inner;
    P.em_k += 1;
    next_emitter(P);
    return false;
  }

  // integrator.rs:402-413: the face ray's result.
  __device__ __forceinline__ void on_emitter(PathState &P, bool has_hit, const Hit &h) const {
    const KernelArgs &ka = kernel_args();
    const DevScene &s = ka.s;
    const lw_params &p = ka.p;
    V3 value = V3{0.0f, 0.0f, 0.0f};
    const float eps = p.misc_f[1];
    if (has_hit && !(f_abs(h.t - P.em_dist) >= 2.0f * eps)) {
      const float4 *hq = s.quads + (size_t)h.quad * 5;
      const V3 c = mul(linear_rgb(lut, h.texel), linear_rgb(lut, f_bits(hq[1].w)));
      value = c * P.em_scale;
    }
    if (face_done(P, value)) next_face(P);
  }

  // integrator.rs:689-693
  __device__ __forceinline__ void add_emitters(PathState &P, V3 e) const {
    P.radiance = P.radiance + mul(P.ba, e);
    finish_bounce(P);
  }
};

}  // namespace lw
