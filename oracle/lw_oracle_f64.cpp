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

}  // extern "C"
