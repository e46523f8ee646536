// ORACLE -- test infrastructure only (see lw_oracle_impl.hpp). The same restatement with
// every operation in double and the transcendental functions libm's double ones (LWO_F64):
// the render that states the rounding error of the f32 path, the reference's CPU backend
// and the device alike. The random draws, constants and branches are the reference's; a
// path can still part from the f32 one where a comparison lands the other way.

#define LWO_NS lwo64
#define LWO_COUNTING 0
#define LWO_F64 1
#include "lw_oracle_impl.hpp"

using namespace lwo64;

extern "C" {

// CpuBackend::render_hq (render/cpu/mod.rs:39-87) in f64: w * h * 3 doubles.
int lwo64_render_hq(const View *view, uint32_t spp, unsigned threads, double *out_rgb) {
  if (threads == 0) threads = 1;
  try {
    render_hq(*view, spp, threads, out_rgb, nullptr);
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}

// render_sample (integrator.rs:539-709) in f64 for n (px, py, sample) triples.
int lwo64_render_samples(const View *view, size_t n, const uint32_t *px, const uint32_t *py,
                         const uint32_t *sample, double *out_rgb) {
  try {
    for (size_t i = 0; i < n; i++) {
      const V3 c = render_sample(*view, px[i], py[i], sample[i]);
      out_rgb[3 * i] = c.x;
      out_rgb[3 * i + 1] = c.y;
      out_rgb[3 * i + 2] = c.z;
    }
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}

// render_hq in f64 on the listed rows only: nrows * w * 3 doubles, row by row in the list's
// order, each pixel the f64 mean of its finite samples (render/cpu/mod.rs:66-79). For
// stating the error of a large frame on a subset of its rows.
int lwo64_render_rows(const View *view, uint32_t spp_in, unsigned threads, const uint32_t *rows,
                      size_t nrows, double *out_rgb) {
  if (threads == 0) threads = 1;
  try {
    Params params = *view->params;
    params.frame[2] = 0;
    View v = *view;
    v.params = &params;
    const uint32_t w = params.frame[0], h = params.frame[1];
    const uint32_t spp = spp_in > 0 ? spp_in : 1;
    for (size_t r = 0; r < nrows; r++)
      if (rows[r] >= h) return -1;
    std::atomic<size_t> next{0};
    std::vector<std::string> errors(threads);
    auto worker = [&](unsigned tid) {
      try {
        for (;;) {
          const size_t r = next.fetch_add(1);
          if (r >= nrows) break;
          for (uint32_t x = 0; x < w; x++) {
            double acc[3] = {0.0, 0.0, 0.0};
            for (uint32_t sample = 0; sample < spp; sample++) {
              const V3 c = render_sample(v, x, rows[r], sample);
              if (finite3(c)) {
                acc[0] += c.x;
                acc[1] += c.y;
                acc[2] += c.z;
              }
            }
            double *px = out_rgb + (r * w + x) * 3;
            for (int k = 0; k < 3; k++) px[k] = acc[k] / (double)spp;
          }
        }
      } catch (const std::exception &e) {
        errors[tid] = e.what();
        next.store(nrows);
      }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < threads; t++) pool.emplace_back(worker, t);
    worker(0);
    for (auto &t : pool) t.join();
    for (auto &e : errors)
      if (!e.empty()) return -1;
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}

}  // extern "C"
