// ORACLE -- test infrastructure only (see lw_oracle_impl.hpp). The same restatement
// compiled with the algorithmic-byte counters on, in its own namespace, so that the
// uncounted build timed as the CPU baseline carries no counting overhead.

#define LWO_NS lwo_counted
#define LWO_COUNTING 1
#include "lw_oracle_impl.hpp"

using namespace lwo_counted;

extern "C" {

// CpuBackend::render_hq (render/cpu/mod.rs:39-87) with the scene-buffer bytes, DDA
// steps, quads tested, trace calls, samples and branch counts summed into *counters, over
// the device's walk (device_walk != 0: it stops once every brick left toward the ray's
// octant is EMPTY) or the reference's whole walk (0).
int lwo_render_hq_counted(const View *view, uint32_t spp, unsigned threads, float *out_rgb,
                          Counters *counters, int device_walk) {
  if (threads == 0) threads = 1;
  try {
    *counters = Counters{};
    render_hq(*view, spp, threads, out_rgb, counters, device_walk != 0);
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}

// The per-brick octant bits the counted walk stops by (tl_octant_boxes), into out (one byte
// per brick).
int lwo_octant_boxes(const View *view, uint8_t *out) {
  try {
    const std::vector<uint8_t> boxes = octant_boxes(*view);
    std::copy(boxes.begin(), boxes.end(), out);
    return 0;
  } catch (const std::exception &) {
    return -1;
  }
}

}  // extern "C"
