// ORACLE -- test infrastructure only (see lw_oracle_impl.hpp). The extern "C" surface
// that tests/ (through ctypes), __graft_entry__.smoke() and bench.py's cpu_baseline leg
// call. Each function names the reference function it restates.

#define LWO_NS lwo
#define LWO_COUNTING 0
#include "lw_oracle_impl.hpp"
#include "lw_oracle_diff.hpp"
#include "lw_oracle_post.hpp"
#include "lw_oracle_build.hpp"
#include "lw_oracle_sky.hpp"
#include "lw_oracle_quick.hpp"

#include <algorithm>
#include <array>
#include <cstdio>

using namespace lwo;

namespace {
thread_local std::string g_err;

template <typename F>
int guarded(F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}
}  // namespace

extern "C" {

const char *lwo_last_error(void) { return g_err.c_str(); }

uint32_t lwo_sizeof(int which) {
  switch (which) {
    case 0: return sizeof(Params);
    case 1: return sizeof(Quad);
    case 2: return sizeof(Model);
    case 3: return sizeof(Emitter);
    case 4: return sizeof(HitOut);
    default: return 0;
  }
}

// rng.rs:5-9
uint32_t lwo_pcg_hash(uint32_t v) { return pcg_hash(v); }
// rng.rs:23-27
uint32_t lwo_rng_new(uint32_t pixel, uint32_t sample, uint32_t seed) {
  return Rng::make(pixel, sample, seed).state;
}
// rng.rs:31-34: n draws from state; returns the new state.
uint32_t lwo_rng_draws(uint32_t state, uint32_t n, float *out) {
  Rng r{state};
  for (uint32_t i = 0; i < n; i++) out[i] = r.next();
  return r.state;
}

// math.rs elementwise over arrays (n values).
void lwo_sin_cos_turns(const float *u, size_t n, float *s, float *c) {
  for (size_t i = 0; i < n; i++) sin_cos_turns(u[i], s[i], c[i]);
}
void lwo_acos(const float *x, size_t n, float *out) {
  for (size_t i = 0; i < n; i++) out[i] = f_acos(x[i]);
}
void lwo_atan2(const float *y, const float *x, size_t n, float *out) {
  for (size_t i = 0; i < n; i++) out[i] = f_atan2(y[i], x[i]);
}
void lwo_ln(const float *x, size_t n, float *out) {
  for (size_t i = 0; i < n; i++) out[i] = f_ln(x[i]);
}
void lwo_gamma_to_linear(const float *x, size_t n, float *out) {
  for (size_t i = 0; i < n; i++) out[i] = gamma_to_linear(x[i]);
}
void lwo_expf(const float *x, size_t n, float *out) {
  for (size_t i = 0; i < n; i++) out[i] = f_exp(x[i]);
}
// math.rs:347-356
void lwo_onb(const float *n3, float *t3, float *b3) {
  V3 t, b;
  onb(v3(n3[0], n3[1], n3[2]), t, b);
  t3[0] = t.x; t3[1] = t.y; t3[2] = t.z;
  b3[0] = b.x; b3[1] = b.y; b3[2] = b.z;
}

// trace.rs:137-290 for n rays.
int lwo_trace(const View *view, size_t n, const float *o, const float *d, const float *tmin,
              const float *tmax, HitOut *out) {
  return guarded([&] {
    for (size_t i = 0; i < n; i++) {
      Hit h;
      HitOut r{};
      bool hit = trace(*view, v3(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                       v3(d[3 * i], d[3 * i + 1], d[3 * i + 2]), tmin[i], tmax[i], h, r.cell);
      if (hit) {
        r.t = h.t;
        r.quad = h.quad;
        r.texel = h.texel;
        r.s = h.s;
        r.tt = h.tt;
        r.hit = 1;
      } else {
        r.t = INFINITY;
        r.quad = UINT32_MAX;
      }
      out[i] = r;
    }
  });
}

// trace.rs:113-116
int lwo_is_see_through(const View *view, uint32_t qi) { return is_see_through(*view, qi) ? 1 : 0; }

// trace.rs:120-131
int lwo_see_through_hit(const View *view, const float *o, const float *d, float tmin, HitOut *out) {
  return guarded([&] {
    Hit h;
    HitOut r{};
    if (see_through_hit(*view, v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), tmin, h)) {
      r.t = h.t;
      r.quad = h.quad;
      r.texel = h.texel;
      r.s = h.s;
      r.tt = h.tt;
      r.hit = 1;
    } else {
      r.quad = UINT32_MAX;
    }
    *out = r;
  });
}

// integrator.rs:539-709 for n (px, py, sample) triples, on `threads` threads.
int lwo_render_samples(const View *view, size_t n, const uint32_t *px, const uint32_t *py,
                       const uint32_t *sample, float *out, unsigned threads) {
  if (threads == 0) threads = 1;
  std::vector<std::string> errors(threads);
  std::atomic<size_t> next{0};
  auto worker = [&](unsigned tid) {
    try {
      for (;;) {
        size_t i0 = next.fetch_add(256);
        if (i0 >= n) break;
        size_t i1 = i0 + 256 < n ? i0 + 256 : n;
        for (size_t i = i0; i < i1; i++) {
          V3 c = render_sample(*view, px[i], py[i], sample[i]);
          out[3 * i] = c.x;
          out[3 * i + 1] = c.y;
          out[3 * i + 2] = c.z;
        }
      }
    } catch (const std::exception &e) {
      errors[tid] = e.what();
      next.store(n);
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < threads; t++) pool.emplace_back(worker, t);
  worker(0);
  for (auto &t : pool) t.join();
  for (auto &e : errors)
    if (!e.empty()) {
      g_err = e;
      return -1;
    }
  return 0;
}

// CpuBackend::render_hq (render/cpu/mod.rs:39-87).
int lwo_render_hq(const View *view, uint32_t spp, unsigned threads, float *out_rgb) {
  if (threads == 0) threads = 1;
  return guarded([&] { render_hq(*view, spp, threads, out_rgb, nullptr); });
}

// integrator.rs:44-74
int lwo_preetham_radiance(const View *view, const float *d, size_t n, float *out) {
  return guarded([&] {
    for (size_t i = 0; i < n; i++) {
      V3 c = preetham_radiance(*view, v3(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
      out[3 * i] = c.x;
      out[3 * i + 1] = c.y;
      out[3 * i + 2] = c.z;
    }
  });
}
// integrator.rs:79-104
void lwo_rgb_to_hsl(const float *rgb, size_t n, float *out) {
  for (size_t i = 0; i < n; i++) rgb_to_hsl(v3(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]), out + 3 * i);
}
// integrator.rs:107-127
void lwo_rgb_from_hsl(const float *hsl, size_t n, float *out) {
  for (size_t i = 0; i < n; i++) {
    V3 c = rgb_from_hsl(hsl[3 * i], hsl[3 * i + 1], hsl[3 * i + 2]);
    out[3 * i] = c.x;
    out[3 * i + 1] = c.y;
    out[3 * i + 2] = c.z;
  }
}
// integrator.rs:133-148
int lwo_build_sky_cache(const View *view, uint32_t res, float *out, size_t out_len) {
  return guarded([&] { build_sky_cache(*view, res, out, out_len); });
}

// integrator.rs:158-188 (which=0), 225-231 (1), 237-243 (2), 252-277 (3).
int lwo_sky(const View *view, int which, const float *d, size_t n, float *out) {
  return guarded([&] {
    for (size_t i = 0; i < n; i++) {
      V3 dd = v3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
      V3 c = which == 0 ? sky_radiance(*view, dd)
                        : which == 1 ? sky_lens(*view, dd)
                                     : which == 2 ? sky_light(*view, dd) : sun_texture(*view, dd);
      out[3 * i] = c.x;
      out[3 * i + 1] = c.y;
      out[3 * i + 2] = c.z;
    }
  });
}

// integrator.rs:357-373
int lwo_emitter_cell(const View *view, const float *p, uint32_t *start_count) {
  return guarded([&] { emitter_cell(*view, v3(p[0], p[1], p[2]), start_count[0], start_count[1]); });
}

// integrator.rs:441-469: `draws` successive calls from one RNG stream; returns the y of
// each in out (n floats * 3).
int lwo_sample_emitters(const View *view, const float *pos, const float *origin, const float *n3,
                        uint32_t rng_state, size_t draws, float *out) {
  return guarded([&] {
    Rng rng{rng_state};
    V3 p = v3(pos[0], pos[1], pos[2]), o = v3(origin[0], origin[1], origin[2]),
       n = v3(n3[0], n3[1], n3[2]);
    for (size_t i = 0; i < draws; i++) {
      V3 c = sample_emitters(*view, p, o, n, rng);
      out[3 * i] = c.x;
      out[3 * i + 1] = c.y;
      out[3 * i + 2] = c.z;
    }
  });
}

// integrator.rs:477-505: returns (d, o, in_water) and the RNG state after the call.
uint32_t lwo_water_interface(const Params *p, const float *pos, const float *d, const float *n,
                             int front, int in_water, float eps, uint32_t rng_state,
                             float *d_out, float *o_out, int *w_out) {
  Rng rng{rng_state};
  V3 dd, oo;
  bool w;
  water_interface(*p, v3(pos[0], pos[1], pos[2]), v3(d[0], d[1], d[2]), v3(n[0], n[1], n[2]),
                  front != 0, in_water != 0, eps, rng, dd, oo, w);
  d_out[0] = dd.x; d_out[1] = dd.y; d_out[2] = dd.z;
  o_out[0] = oo.x; o_out[1] = oo.y; o_out[2] = oo.z;
  *w_out = w ? 1 : 0;
  return rng.state;
}


// shadow (integrator.rs:316-346): the transmittance from o along d up to tmax, for n rays.
int lwo_shadow(const View *view, size_t n, const float *o, const float *d, const float *tmax, float *out) {
  return guarded([&] {
    for (size_t i = 0; i < n; i++) {
      const V3 t = shadow(*view, v3(o[3 * i], o[3 * i + 1], o[3 * i + 2]), v3(d[3 * i], d[3 * i + 1], d[3 * i + 2]),
                          tmax[i]);
      out[3 * i] = t.x;
      out[3 * i + 1] = t.y;
      out[3 * i + 2] = t.z;
    }
  });
}

// sun_texture (integrator.rs:252-277) along n directions.
int lwo_sun_texture(const View *view, size_t n, const float *d, float *out) {
  return guarded([&] {
    for (size_t i = 0; i < n; i++) {
      const V3 c = sun_texture(*view, v3(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
      out[3 * i] = c.x;
      out[3 * i + 1] = c.y;
      out[3 * i + 2] = c.z;
    }
  });
}

// Differentiable rendering's CPU reference (lw_oracle_diff.hpp): the image under the
// parameters of a lw_diff_desc, and the gradient of sum(dldi * image).
int lwo_render_diff(const View *view, uint32_t spp, const DiffDesc *diff, unsigned threads,
                    float *out_rgb) {
  if (threads == 0) threads = 1;
  return guarded([&] { render_diff(*view, spp, *diff, threads, out_rgb); });
}

// sum(weights * image) of render_diff with the image in f64 (finite differences).
int lwo_render_diff_dot(const View *view, uint32_t spp, const DiffDesc *diff, const float *weights,
                        unsigned threads, double *out) {
  if (threads == 0) threads = 1;
  return guarded([&] {
    std::vector<float> img((size_t)view->params->frame[0] * view->params->frame[1] * 3);
    render_diff(*view, spp, *diff, threads, img.data(), weights, out);
  });
}

// d sky_radiance(d) / d(sky parameter k) for n directions (lw_oracle_diff.hpp).
int lwo_sky_tangent(const View *view, const DiffDesc *diff, int k, const float *d, size_t n,
                    float *out) {
  return guarded([&] {
    if (k < 0 || k > 2) throw OracleError("k is 0, 1 or 2");
    check_diff(*view, *diff);
    const DiffCtx C{*diff, view->quads_len, cache_len_of(*view), nullptr};
    for (size_t i = 0; i < n; i++) {
      V3 c = sky_radiance_tangent(*view, C, k, v3(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
      out[3 * i] = c.x;
      out[3 * i + 1] = c.y;
      out[3 * i + 2] = c.z;
    }
  });
}

int lwo_adjoint(const View *view, uint32_t spp, const DiffDesc *diff, const float *dldi,
                unsigned threads, double *grads) {
  if (threads == 0) threads = 1;
  return guarded([&] { adjoint(*view, spp, *diff, dldi, threads, grads); });
}


// ---- post-processing (render/post.rs), lw_oracle_post.hpp ----

// tonemap (post.rs:148-170); op: 0 Gamma, 1 Aces, 2 Filmic, 3 FilmicMidGrey.
void lwo_tonemap(const float *hdr, size_t n, float exposure, uint32_t op, uint32_t *out) {
  lwo_post::tonemap(hdr, n, exposure, op, out);
}

// tonemap_mix (post.rs:176-215)
void lwo_tonemap_mix(const float *hdr, size_t n, float exposure, uint32_t a, uint32_t b,
                     float weight, uint32_t *out) {
  lwo_post::tonemap_mix(hdr, n, exposure, a, b, weight, out);
}

// downsample (post.rs:218-247); out holds (w / k) * (h / k) words (w * h when k <= 1).
void lwo_downsample(const uint32_t *px, uint32_t w, uint32_t h, uint32_t k, uint32_t *out) {
  auto v = lwo_post::downsample(px, w, h, k);
  std::memcpy(out, v.data(), v.size() * 4);
}

// The Hable curve of an op (post.rs:37-138): a, b, c, d, e, f, w, white_scale; and the
// game curve's gain mid_grey_gain(0.18) and before_gamma(x) for x.
void lwo_hable(uint32_t op, float out[8], double *gain, double x, double *before) {
  const lwo_post::HableCurve c = lwo_post::curve_of(op);
  const float v[8] = {c.a, c.b, c.c, c.d, c.e, c.f, c.w, c.white_scale()};
  std::memcpy(out, v, sizeof(v));
  *gain = lwo_post::HABLE_FILMIC_WORLDS.mid_grey_gain(lwo_post::MID_GREY);
  *before = c.before_gamma(x);
}

// encode (post.rs:141-144) elementwise.
void lwo_encode(const float *v, size_t n, float *out) {
  for (size_t i = 0; i < n; i++) out[i] = lwo_post::encode(v[i]);
}

// The PFM file of an image (output.rs:19-30); returns its size, writing it when cap allows.
size_t lwo_pfm(const float *hdr, uint32_t w, uint32_t h, uint8_t *out, size_t cap) {
  std::string s = lwo_post::pfm(hdr, w, h);
  if (out && cap >= s.size()) std::memcpy(out, s.data(), s.size());
  return s.size();
}

// Every f32 v in [0, 1] (and 1.5, +inf): the count of thresholds <= v against post.rs's
// byte (encode(v) * 255 + 0.5) as u32, on `threads` threads. Returns the number of
// mismatches and the first mismatching input in *first.
uint64_t lwo_check_encode_thresholds(const float *thr, unsigned threads, uint32_t *first) {
  const uint32_t one = 0x3f800000u;
  if (threads == 0) threads = 1;
  for (int k = 1; k < 255; k++)
    if (!(thr[k - 1] <= thr[k])) {
      *first = (uint32_t)k;
      return ~0ull;  // not sorted
    }
  std::vector<uint64_t> bad(threads, 0);
  std::vector<uint32_t> fst(threads, 0xFFFFFFFFu);
  auto work = [&](unsigned t) {
    const uint32_t lo = (uint32_t)((uint64_t)(one + 1) * t / threads);
    const uint32_t hi = (uint32_t)((uint64_t)(one + 1) * (t + 1) / threads);
    for (uint32_t u = lo; u < hi; u++) {
      float v;
      std::memcpy(&v, &u, 4);
      uint32_t want = lwo_post::as_u32_sat(lwo_post::encode(v) * 255.0f + 0.5f);
      uint32_t got = (uint32_t)(std::upper_bound(thr, thr + 255, v) - thr);  // sorted: checked
      if (got != want) {
        if (!bad[t]) fst[t] = u;
        bad[t]++;
      }
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < threads; t++) pool.emplace_back(work, t);
  work(0);
  for (auto &t : pool) t.join();
  uint64_t total = 0;
  *first = 0xFFFFFFFFu;
  for (unsigned t = 0; t < threads; t++) {
    total += bad[t];
    if (bad[t] && fst[t] < *first) *first = fst[t];
  }
  return total;
}

// libm against its function in f64 rounded once to f32, over every f32 input in a range: powf(x,
// y) over [0, 1] (y = 1/2.2f, the display encoding, or 2.2f, the texel decoding) and log2f over
// [1, 65536] (the quick mode's mip level). A deviation is an input where the two differ. The
// distance of the f64 value p from the f32 rounding midpoint on p's side of round(p) (2^-150
// when p rounds to 0) is taken relative to |p| for powf and absolute for log2f; the input is in
// the band when that distance is below `band` (2^-32 relative, 2^-28 absolute). stats: [0]
// deviations, [1] deviations outside the band, [2] inputs in the band, [3] deviations that are
// not a neighbour of round(p); *max_dist: the largest distance among the deviations. With
// `band` non-null (room for `cap`), the band inputs are written there in ascending order.
extern "C++" {
static bool band_of(double p, double band, bool relative, double &dist) {
  const double ap = std::fabs(p);
  const float c = (float)ap;
  dist = relative ? 1.0 : 1e300;
  if (c == 0.0f && ap > 0.0) {
    dist = std::fabs(ap - 0x1p-150) / (relative ? ap : 1.0);
    return dist < band;
  }
  if (!(c > 0.0f && c <= 3.4028234e38f)) return false;
  uint32_t cb;
  std::memcpy(&cb, &c, 4);
  const uint32_t nbb = (double)c <= ap ? cb + 1u : cb - 1u;
  float nb;
  std::memcpy(&nb, &nbb, 4);
  const double mid = ((double)c + (double)nb) * 0.5;
  dist = std::fabs(ap - mid) / (relative ? ap : 1.0);
  return dist < band;
}

template <typename Libm, typename Exact>
static void scan_band(uint32_t u0, uint32_t u1, Libm libm, Exact exact, double width, bool relative,
                      unsigned threads, uint64_t stats[4], double *max_rel, uint32_t *band, uint64_t cap) {
  if (threads == 0) threads = 1;
  std::vector<std::array<uint64_t, 4>> st(threads, std::array<uint64_t, 4>{0, 0, 0, 0});
  std::vector<double> mr(threads, 0.0);
  std::vector<std::vector<uint32_t>> found(threads);
  const uint64_t n = (uint64_t)u1 - u0 + 1;
  auto work = [&](unsigned t) {
    const uint64_t lo = u0 + n * t / threads, hi = u0 + n * (t + 1) / threads;
    for (uint64_t w = lo; w < hi; w++) {
      const uint32_t u = (uint32_t)w;
      float x;
      std::memcpy(&x, &u, 4);
      const double p = exact(x);
      double rel;
      const bool in = band_of(p, width, relative, rel);
      if (in) {
        if (band) found[t].push_back(u);
        st[t][2]++;
      }
      const float g = libm(x);
      const float c = (float)p;
      uint32_t gb, cb;
      std::memcpy(&gb, &g, 4);
      std::memcpy(&cb, &c, 4);
      if (gb != cb) {
        st[t][0]++;
        if (!in) st[t][1]++;
        if (gb != cb + 1u && gb != cb - 1u) st[t][3]++;
        mr[t] = std::max(mr[t], rel);
      }
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < threads; t++) pool.emplace_back(work, t);
  work(0);
  for (auto &t : pool) t.join();
  for (int k = 0; k < 4; k++) stats[k] = 0;
  *max_rel = 0.0;
  uint64_t at = 0;
  for (unsigned t = 0; t < threads; t++) {
    for (int k = 0; k < 4; k++) stats[k] += st[t][k];
    *max_rel = std::max(*max_rel, mr[t]);
    for (uint32_t u : found[t])
      if (at < cap) band[at++] = u;
  }
}
}  // extern "C++"

void lwo_powf_band(float y, unsigned threads, uint64_t stats[4], double *max_rel, uint32_t *band,
                   uint64_t cap) {
  scan_band(0u, 0x3f800000u, [y](float x) { return std::pow(x, y); },
            [y](float x) { return std::pow((double)x, (double)y); }, 0x1p-32, true, threads, stats, max_rel, band,
            cap);
}

void lwo_log2f_band(unsigned threads, uint64_t stats[4], double *max_rel, uint32_t *band, uint64_t cap) {
  scan_band(0x3f800000u, 0x47800000u, [](float x) { return std::log2(x); },
            [](float x) { return std::log2((double)x); }, 0x1p-28, false, threads, stats, max_rel, band, cap);
}

// map_direction (sky.rs:856-860) and blend_chunky_sky (sky.rs:836-851) with the kernel's
// sky_radiance through the view's sky cache as Chunky's sky (request.rs:891-911): maps holds
// the lens's then the light's map, sizes = view params' sky_map.
int lwo_blend_chunky_sky(const View *view, float weight, const float *sizes, float *maps) {
  return guarded([&] {
    const float TAU = 6.28318530717958647692f, PI = 3.14159265358979323846f,
                FRAC_PI_2 = 1.57079632679489661923f;
    const uint32_t dims[2][2] = {{(uint32_t)sizes[0], (uint32_t)sizes[1]},
                                 {(uint32_t)sizes[2], (uint32_t)sizes[3]}};
    float *buf = maps;
    for (auto &wh : dims) {
      const uint32_t w = wh[0], h = wh[1];
      for (uint32_t k = 0; k < w * h; k++) {
        const float az = ((float)(k % w) + 0.5f) / (float)w * TAU;
        const float el = ((float)(k / w) + 0.5f) / (float)h * PI - FRAC_PI_2;
        const V3 d = v3(std::cos(el) * std::cos(az), std::sin(el), std::cos(el) * std::sin(az));
        const V3 c = sky_radiance(*view, d);
        float *t = buf + (size_t)k * 3;
        t[0] = weight * c.x + (1.0f - weight) * t[0];
        t[1] = weight * c.y + (1.0f - weight) * t[1];
        t[2] = weight * c.z + (1.0f - weight) * t[2];
      }
      buf += (size_t)w * h * 3;
    }
  });
}

// ---- the game's sky (render/sky.rs) ----

// java.util.Random (sky.rs:72-98): n values of next(bits) from seed.
void lwo_java_random(int64_t seed, int bits, size_t n, int32_t *out) {
  lwo_gsky::Random r(seed);
  for (size_t i = 0; i < n; i++) out[i] = r.next(bits);
}

// stars (sky.rs:109-137) as 9 floats each (center, a1, a2); returns the count.
long lwo_stars(float *out, size_t cap) {
  const auto st = lwo_gsky::stars();
  if (out && cap >= st.size() * 9)
    for (size_t i = 0; i < st.size(); i++) {
      const lwo_gsky::V v[3] = {st[i].center, st[i].a1, st[i].a2};
      for (int k = 0; k < 3; k++) {
        out[9 * i + 3 * k] = v[k].x;
        out[9 * i + 3 * k + 1] = v[k].y;
        out[9 * i + 3 * k + 2] = v[k].z;
      }
    }
  return (long)st.size();
}

// star_table (sky.rs:152-222); returns the word count.
long lwo_star_table(uint32_t *out, size_t cap) {
  static const std::vector<uint32_t> t = lwo_gsky::star_table();
  if (out && cap >= t.size()) std::copy(t.begin(), t.end(), out);
  return (long)t.size();
}

// QuickSky::new (sky.rs:320-373); cam = forward, right, up, tan(fov/2), aspect.
void lwo_quick_sky_new(const float *cam, uint32_t w, uint32_t h, const lwo_gsky::Description *d,
                       lwo_gsky::Quick *out) {
  *out = lwo_gsky::quick_new(cam, w, h, *d);
}

void lwo_celestial_direction(float deg, float *out) {
  const lwo_gsky::V v = lwo_gsky::celestial_direction(deg);
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
}

void lwo_lightmap_sky(float factor, const float *colour, float brightness, float *out) {
  const lwo_gsky::V v = lwo_gsky::lightmap_sky(factor, {colour[0], colour[1], colour[2]}, brightness);
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
}

// QuickSky::background (sky.rs:509-532) along n directions.
void lwo_quick_background(const lwo_gsky::Quick *q, const float *d, size_t n, const uint32_t *atlas,
                          uint32_t atlas_w, float *out) {
  static const std::vector<uint32_t> t = lwo_gsky::star_table();
  for (size_t i = 0; i < n; i++) {
    const lwo_gsky::V c = lwo_gsky::background(*q, {d[3 * i], d[3 * i + 1], d[3 * i + 2]}, atlas, atlas_w, t);
    out[3 * i] = c.x;
    out[3 * i + 1] = c.y;
    out[3 * i + 2] = c.z;
  }
}

// game_sky_maps (sky.rs:906-953): out gets both maps; info light[3], mean[3], scale, veil[3].
void lwo_game_sky_maps(const lwo_gsky::Quick *lens, const lwo_gsky::Quick *light, float factor,
                       const float *colour, const uint32_t *atlas, uint32_t atlas_w, unsigned threads, float *out,
                       float *info) {
  static const std::vector<uint32_t> t = lwo_gsky::star_table();
  lwo_gsky::game_sky_maps(*lens, *light, factor, {colour[0], colour[1], colour[2]}, atlas, atlas_w, t, out, info,
                          threads ? threads : 1);
}

// rasterize (raster.rs:168-372): a quick frame's packed RGBA8 pixels, row major, with the
// client's stars, the tiles on `threads` threads. lists: the opaque, translucent and
// see-through index lists, list_len their lengths in indices.
int lwo_rasterize(const lwo_quick::Vertex *verts, size_t n_verts, const uint32_t *const *lists,
                  const size_t *list_len, const uint32_t *atlas, const uint32_t *atlas_size, const uint32_t *mips,
                  uint32_t levels, uint32_t mip_rows, uint32_t width, uint32_t height, const float *view_proj,
                  const lwo_gsky::Quick *sky, unsigned threads, uint32_t *out) {
  return guarded([&] {
    static const std::vector<uint32_t> table = lwo_gsky::star_table();
    for (int k = 0; k < 3; k++)
      for (size_t i = 0; i < list_len[k] / 3 * 3; i++)
        if (lists[k][i] >= n_verts) throw std::runtime_error("vertex index out of range");
    lwo_quick::Frame f;
    f.verts = verts;
    f.n_verts = n_verts;
    for (int k = 0; k < 3; k++) {
      f.lists[k] = lists[k];
      f.list_len[k] = list_len[k];
    }
    f.atlas = atlas;
    f.aw = atlas_size[0];
    f.ah = atlas_size[1];
    f.mips = mips;
    f.levels = levels;
    f.mip_rows = mip_rows;
    f.width = width;
    f.height = height;
    std::memcpy(f.view_proj, view_proj, sizeof(f.view_proj));
    f.sky = sky;
    const std::vector<uint32_t> px = lwo_quick::rasterize(f, table, threads ? threads : 1);
    std::copy(px.begin(), px.end(), out);
  });
}

// AtlasView (texture.rs:46-145) at n points q = (u, v, du.x, du.y, dv.x, dv.y): mode 0 sample,
// 1 sample_block, 2 sample_grad. out: n RGBA in [0, 1].
void lwo_atlas_sample(const uint32_t *level0, const uint32_t *mips, const uint32_t *size, uint32_t levels,
                      uint32_t mip_rows, int mode, const float *q, size_t n, float *out) {
  const lwo_quick::AtlasView v{level0, mips, size[0], size[1], levels, mip_rows};
  for (size_t i = 0; i < n; i++) {
    const float *a = q + 6 * i;
    lwo_quick::V4 c;
    if (mode == 2) c = v.sample_grad(a[0], a[1], a[2], a[3], a[4], a[5]);
    else if (mode == 1 || v.is_mipmapped(a[1])) c = v.sample_block(a[0], a[1], a[2], a[3], a[4], a[5]);
    else c = v.nearest(a[0], a[1]);
    out[4 * i] = c.x;
    out[4 * i + 1] = c.y;
    out[4 * i + 2] = c.z;
    out[4 * i + 3] = c.w;
  }
}

// build_hq's layout (voxel.rs:305-491, emitter_list :542-607) of an interned volume. sizes:
// grid, emitters, emitter grid lengths; meta: offsets[5], see_through[2], emitter_count;
// stats: bricks, uniform, buried, overlay bricks. The buffers are fetched with
// lwo_build_fetch (the last build on this thread).
static thread_local lwo_build::Built g_built;

int lwo_build_layout(const uint32_t *dims, const uint16_t *voxels, const uint8_t *brick_occludes,
                     const lwo_build::Model *models, size_t models_len, const uint8_t *flags,
                     size_t quads_len, const lwo_build::Quad *entity, size_t n_entity,
                     uint32_t see_through_count, uint64_t *sizes, uint32_t *meta, uint64_t *stats) {
  return guarded([&] {
    lwo_build::Volume v{{dims[0], dims[1], dims[2]}, voxels, brick_occludes, models, models_len, flags,
                        quads_len, entity, n_entity, see_through_count};
    g_built = lwo_build::build(v);
    sizes[0] = g_built.grid.size();
    sizes[1] = g_built.emitters.size();
    sizes[2] = g_built.emitter_grid.size();
    for (int k = 0; k < 5; k++) meta[k] = g_built.offsets[k];
    meta[5] = g_built.see_through[0];
    meta[6] = g_built.see_through[1];
    meta[7] = g_built.emitter_count;
    stats[0] = g_built.bricks;
    stats[1] = g_built.uniform_bricks;
    stats[2] = g_built.buried_bricks;
    stats[3] = g_built.overlay_bricks;
  });
}

void lwo_build_fetch(uint32_t *grid, lwo_build::Emitter *emitters, uint32_t *egrid) {
  std::copy(g_built.grid.begin(), g_built.grid.end(), grid);
  std::copy(g_built.emitters.begin(), g_built.emitters.end(), emitters);
  std::copy(g_built.emitter_grid.begin(), g_built.emitter_grid.end(), egrid);
}

}  // extern "C"
